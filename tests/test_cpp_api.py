"""Host-side checks of the C++ host API (include/flsgpu.hpp), no GPU: the
combinations it accepts match the reference's generated bindings one for
one, and the reference-style command line rejects bad arguments the way the
reference does (std::invalid_argument, reported as exit status 100)."""
import os
import subprocess

import pytest

HERE = os.path.dirname(__file__)
BUILD = os.path.join(HERE, "cpp", "_build")

# Counts of (unpack_n_vectors, unpack_n_values, unpacker, patcher) branches in
# the bindings generated by generate-kernel-bindings.py (loops :284-396,
# filters :241-268): 17 regular unpackers x n_vec {1, 4}, plus old-fls
# (n_vec 1, n_val 32, uint32/float only) and switch-case (n_vec 1, unsigned,
# not compute); ALP drops the dummy unpacker; patchers per encoding
# {none} / {dummy, stateless, stateful} / {naive .. prefetch-all-branchless};
# multi-column keeps {old-fls, stateful-register-branchless-2,
# stateful-branchless} x {none, stateful, naive, naive-branchless,
# prefetch-all, prefetch-all-branchless}.
EXPECTED = {
    ("decompress", "BP", "uint32_t"): 36, ("decompress", "BP", "uint64_t"): 35,
    ("decompress", "FFOR", "uint32_t"): 36, ("decompress", "FFOR", "uint64_t"): 35,
    ("query", "BP", "uint32_t"): 36, ("query", "BP", "uint64_t"): 35,
    ("query", "FFOR", "uint32_t"): 36, ("query", "FFOR", "uint64_t"): 35,
    ("decompress", "ALP", "float"): 99, ("decompress", "ALP", "double"): 96,
    ("decompress", "ALPExtended", "float"): 165, ("decompress", "ALPExtended", "double"): 160,
    ("query", "ALP", "float"): 99, ("query", "ALP", "double"): 96,
    ("query", "ALPExtended", "float"): 165, ("query", "ALPExtended", "double"): 160,
    ("compute", "FFOR", "uint32_t"): 35, ("compute", "FFOR", "uint64_t"): 34,
    ("compute", "BP", "uint32_t"): 0, ("compute", "BP", "uint64_t"): 0,
    ("compute", "ALP", "float"): 0, ("compute", "ALP", "double"): 0,
    ("compute", "ALPExtended", "float"): 0, ("compute", "ALPExtended", "double"): 0,
    ("multi", "BP", "uint32_t"): 0, ("multi", "BP", "uint64_t"): 0,
    ("multi", "FFOR", "uint32_t"): 5, ("multi", "FFOR", "uint64_t"): 4,
    ("multi", "ALP", "float"): 5, ("multi", "ALP", "double"): 4,
    ("multi", "ALPExtended", "float"): 20, ("multi", "ALPExtended", "double"): 16,
}


def _exe(name):
    p = os.path.join(BUILD, name)
    if not os.path.exists(p):
        pytest.skip(f"{p} not built (make -C tests/cpp)")
    return p


def test_provided_combinations_match_generated_bindings():
    out = subprocess.run([_exe("api_check")], capture_output=True, text=True, check=True).stdout
    got = {}
    for line in out.splitlines():
        parts = line.split()
        if parts[0] == "enum_errors":
            assert parts[1] == "0"
        else:
            got[tuple(parts[:3])] = int(parts[3])
    assert got == EXPECTED


@pytest.mark.parametrize("args", [
    ["u32"],                                                                  # arg count
    ["u32", "decompress", "1", "1", "no-such", "none", "0", "8", "0", "0", "4", "1", "0"],
    ["x32", "decompress", "1", "1", "stateless", "none", "0", "8", "0", "0", "4", "1", "0"],
    ["u32", "decompress", "1", "1", "stateless", "bogus", "0", "8", "0", "0", "4", "1", "0"],
])
def test_cli_rejects_bad_arguments(args):
    p = subprocess.run([_exe("flsgpu_test")] + args, capture_output=True, text=True, timeout=60)
    assert p.returncode == 100, p.stderr
    assert "invalid_argument" in p.stderr
