"""Encode a data file and decode or query it on the GPU: the reference's
real-data ALP benchmark (src/bench-compressors.cu:24-82 command line,
src/nvcomp/benchmark-compressors.cu:86-178 benchmark_alp,
src/engine/data.cuh:184-233 read_file_as) on this library.

    python -m flsgpu.e2e <f32|f64> <decompression|decompression_query>
                         <ALP|GALP|ALPRD> <data_path> <vector_count> [options]

The file holds raw little-endian values of the data type. As in the
reference, vector_count * 1024 values are used (0: every whole vector in the
file), a file shorter than that is repeated to fill the column, and the
query searches for the value whose bits are 0x4100f9db (f32) or
0x000000004100f9db (f64). Output is the reference's CSV line
    comparison,compressor,data_name,found,n_bytes,execution_time_ms,compression_ratio
where `found` is 1 when the decoded column equals the input bit for bit
(decompression) or when the query value occurs (decompression_query; the
reference prints 0 there), and execution_time_ms is one kernel launch (the
mean over --repeats launches). ALPRD (an extension: the reference has no
ALP-RD decoder) encodes with ALP-RD; ALP and GALP fail, as in the reference,
when sampling selects ALP-RD. --via-file writes the encoded column to a
column file and loads it back before decoding. --json prints a second line
with throughput, roofline fraction and the timings of every stage.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HBM_PEAK_GBS = 8000.0
QUERY_BITS = 0x4100F9DB


def read_file_as(path: str, dtype, n_values: int) -> np.ndarray:
    """data::arrays::read_file_as (src/engine/data.cuh:184-233)."""
    size = os.path.getsize(path)
    item = np.dtype(dtype).itemsize
    if size % item:
        raise ValueError("File size is incorrect, it is not a multiple of the type's size.")
    in_file = size // item
    count = in_file if n_values == 0 else n_values
    count -= count % 1024
    if count == 0 or in_file == 0:
        raise ValueError("the column would hold no whole vector")
    data = np.fromfile(path, dtype=dtype, count=min(in_file, count))
    if len(data) < count:
        reps = -(-count // len(data))
        data = np.tile(data, reps)[:count]
    return np.ascontiguousarray(data)


def query_value(dtype):
    return np.array([QUERY_BITS], dtype=np.uint32 if dtype == np.float32 else np.uint64).view(dtype)[0]


def data_name(path: str) -> str:
    base = os.path.basename(path)
    return base.rsplit(".", 1)[0] if "." in base else base


def run(args) -> dict:
    import torch
    from . import _capi as A
    from . import alp as AL
    from . import colfile as CF

    dtype = np.float32 if args.data_type == "f32" else np.float64
    w = 32 if dtype == np.float32 else 64
    t0 = time.perf_counter()
    x = read_file_as(args.data_path, dtype, args.vector_count * 1024)
    t_read = time.perf_counter() - t0
    A.check(A.lib().flsgpu_init(0))

    t0 = time.perf_counter()
    if args.compressor == "ALPRD":
        host = AL.alprd_encode(x)
        ratio = x.nbytes / host.compressed_bytes
    else:
        host = AL.alp_encode(x)
        if host is None:
            raise RuntimeError("ALP sampling selected ALP_RD for this data "
                               "(alp::encode throws here); use ALPRD")
        ratio = host.compression_ratio(extended=args.compressor == "GALP")
        if args.compressor == "GALP":
            host = AL.galp_convert(host)
    t_encode = time.perf_counter() - t0

    t_file = None
    if args.via_file:
        t0 = time.perf_counter()
        CF.write(args.via_file, host)
        col = CF.load(args.via_file)
        t_file = time.perf_counter() - t0
        dev = None
    else:
        dev = (AL.DeviceAlprdColumn.from_host(host) if args.compressor == "ALPRD"
               else AL.DeviceAlpColumn.from_host(host))
        col = None
    torch.cuda.synchronize()

    out = torch.empty(len(x), dtype=torch.float32 if w == 32 else torch.float64, device="cuda")
    magic = query_value(dtype)
    stream = torch.cuda.current_stream()

    def decode():
        if col is not None:
            col.decompress(out, stream=stream)
        elif args.compressor == "ALPRD":
            AL.alprd_decompress(dev, out, stream=stream)
        else:
            AL.decompress(dev, out, stream=stream)

    def query():
        if col is not None:
            return col.query_any(float(magic), stream=stream)
        return AL.query_any(dev, float(magic), stream=stream)

    op = decode if args.comparison == "decompression" else query
    if args.comparison == "decompression_query" and args.compressor == "ALPRD":
        raise RuntimeError("decompression_query has no ALP-RD kernel")
    op()
    torch.cuda.synchronize()
    a = torch.cuda.Event(enable_timing=True)
    b = torch.cuda.Event(enable_timing=True)
    a.record(stream)
    for _ in range(args.repeats):
        res = op()
    b.record(stream)
    torch.cuda.synchronize()
    ms = a.elapsed_time(b) / args.repeats

    if args.comparison == "decompression":
        got = out.cpu().numpy()
        ut = np.uint32 if w == 32 else np.uint64
        found = int(np.array_equal(got.view(ut), x.view(ut)))
    else:
        found = int(res.item())
    compressed = x.nbytes / ratio
    read_bytes = compressed
    write_bytes = x.nbytes if args.comparison == "decompression" else 0
    return {
        "csv": f"{args.comparison},{args.compressor},{data_name(args.data_path)},{found},"
               f"{x.nbytes},{ms:f},{ratio:f}",
        "found": found,
        "n_values": len(x),
        "n_bytes": x.nbytes,
        "kernel_ms": ms,
        "decompressed_GBps": x.nbytes / ms / 1e6,
        "hbm_frac": (read_bytes + write_bytes) / ms / 1e6 / HBM_PEAK_GBS,
        "compression_ratio": ratio,
        "bits_per_value": 8 * compressed / len(x),
        "read_s": t_read,
        "encode_s": t_encode,
        "file_write_load_s": t_file,
        "query_value_present": bool(np.any(x == magic)),
    }


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("data_type", choices=["f32", "f64"])
    ap.add_argument("comparison", choices=["decompression", "decompression_query"])
    ap.add_argument("compressor", choices=["ALP", "GALP", "ALPRD"])
    ap.add_argument("data_path")
    ap.add_argument("vector_count", type=int)
    ap.add_argument("--via-file", default=None, help="write + load this column file first")
    ap.add_argument("--repeats", type=int, default=10)
    ap.add_argument("--json", action="store_true")
    args = ap.parse_args(argv)
    r = run(args)
    print(r["csv"], flush=True)
    if args.json:
        print(json.dumps({k: v for k, v in r.items() if k != "csv"}), flush=True)
    return 0 if args.comparison == "decompression_query" or r["found"] else 1


if __name__ == "__main__":
    sys.exit(main())
