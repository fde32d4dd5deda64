"""The reference's command-line test matrix (test-fls.sh, test-alp.sh,
test-multi-column.sh: ./bin/test with 13 arguments) run through the C++
host API (include/flsgpu.hpp) by tests/cpp/flsgpu_test, which checks every
decoded value bit for bit against the CPU oracle. Exit status 0 = all runs
exact; 100 = the combination is not provided (std::invalid_argument, as the
reference's generated bindings throw)."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

EXE = os.path.join(os.path.dirname(__file__), "cpp", "_build", "flsgpu_test")
N_VECS = "64"
UNPACKERS = ["stateless", "stateless-branchless", "stateful-cache", "stateful-local-1",
             "stateful-local-2", "stateful-local-4", "stateful-shared-1", "stateful-shared-2",
             "stateful-shared-4", "stateful-register-1", "stateful-register-2",
             "stateful-register-4", "stateful-register-branchless-1",
             "stateful-register-branchless-2", "stateful-register-branchless-4",
             "stateful-branchless"]
ALP_PATCHERS = ["stateless", "stateful", "naive", "naive-branchless", "prefetch-position",
                "prefetch-all", "prefetch-all-branchless"]


def run(*args):
    if not os.path.exists(EXE):
        pytest.fail(f"{EXE} missing: build it with `make -C tests/cpp`")
    argv = [EXE] + [str(a) for a in args]
    print(" ".join(argv[1:]), flush=True)
    p = subprocess.run(argv, capture_output=True, text=True, timeout=120)
    return p.returncode, p.stderr


def case_ids(cases):
    return [" ".join(str(x) for x in c) for c in cases]


# test-fls.sh "Main decompressors": u32/u64 x decompress/query x n_vec 1,4, all widths
FLS_CASES = ([(t, k, nv, u) for t in ("u32", "u64") for k in ("decompress", "query")
              for nv in (1, 4) for u in ("stateful-branchless", "stateless")]
             + [("u32", "decompress", 1, u) for u in UNPACKERS])


@pytest.mark.parametrize("t,kernel,nvec,unpacker", FLS_CASES, ids=case_ids(FLS_CASES))
def test_fls_main(t, kernel, nvec, unpacker):
    w = t[1:]
    rc, err = run(t, kernel, nvec, 1, unpacker, "none", 0, w, 0, 0, N_VECS, 1, 1)
    assert rc == 0, err


@pytest.mark.parametrize("t", ["u32", "u64"])
@pytest.mark.parametrize("kernel", ["decompress", "query"])
def test_fls_dummy_and_special_unpackers(t, kernel):
    w = t[1:]
    # dummy unpacker at full width (test-fls.sh "Dummy decompressor")
    assert run(t, kernel, 4, 1, "dummy", "none", w, w, 0, 0, N_VECS, 1, 1)[0] == 0
    # switch-case: uint, n_vec 1, n_val 1
    assert run(t, kernel, 1, 1, "switch-case", "none", 0, w, 0, 0, N_VECS, 1, 1)[0] == 0
    # old-fls: u32 only, n_vec 1, n_val 32
    rc = run(t, kernel, 1, 32, "old-fls", "none", 0, w, 0, 0, N_VECS, 1, 1)[0]
    assert rc == (0 if t == "u32" else 100)


ALP_CASES = [(t, k, p) for t in ("f32", "f64") for k in ("decompress", "query")
             for p in ALP_PATCHERS]


@pytest.mark.parametrize("t,kernel,patcher", ALP_CASES, ids=case_ids(ALP_CASES))
def test_alp_main(t, kernel, patcher):
    # test-alp.sh "Main decompressors": widths 0..16, 0..10 exceptions per vector
    rc, err = run(t, kernel, 4, 1, "stateful-branchless", patcher, 0, 16, 0, 10, N_VECS, 1, 1)
    assert rc == 0, err


@pytest.mark.parametrize("t", ["f32", "f64"])
def test_alp_dummy_patcher_and_old_fls(t):
    # dummy patcher without exceptions (test-alp.sh "Dummy patcher")
    assert run(t, "decompress", 1, 1, "stateless", "dummy", 0, 16, 0, 0, N_VECS, 1, 1)[0] == 0
    # old-fls: float only
    rc = run(t, "decompress", 1, 32, "old-fls", "prefetch-all", 0, 16, 0, 10, N_VECS, 1, 1)[0]
    assert rc == (0 if t == "f32" else 100)


MULTI_CASES = ([(t, "none") for t in ("u32", "u64")]
               + [(t, p) for t in ("f32", "f64")
                  for p in ("stateful", "naive", "naive-branchless", "prefetch-all",
                            "prefetch-all-branchless")])


@pytest.mark.parametrize("t,patcher", MULTI_CASES, ids=case_ids(MULTI_CASES))
def test_multi_column(t, patcher):
    # test-multi-column.sh: widths 0..8, 20 exceptions per vector
    ec = 0 if t.startswith("u") else 20
    rc, err = run(t, "query-multi-column", 1, 1, "stateful-branchless", patcher, 0, 8, ec, ec,
                  N_VECS, 1, 1)
    assert rc == 0, err


def test_invalid_combinations_throw():
    # not provided by the reference's generated bindings
    assert run("f32", "decompress", 1, 1, "switch-case", "stateful", 0, 8, 0, 0, 4, 1, 0)[0] == 100
    assert run("f64", "decompress", 1, 1, "stateless", "none", 0, 8, 0, 0, 4, 1, 0)[0] == 100
    assert run("u32", "decompress", 2, 1, "stateless", "none", 0, 8, 0, 0, 4, 1, 0)[0] == 100
    assert run("f32", "query-multi-column", 1, 1, "stateful-branchless", "dummy", 0, 8, 0, 0,
               4, 1, 0)[0] == 100
    assert run("u32", "decompress", 1, 1, "no-such-unpacker", "none", 0, 8, 0, 0, 4, 1, 0)[0] == 100
