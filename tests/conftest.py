import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TESTS = os.path.dirname(os.path.abspath(__file__))
for p in (ROOT, TESTS):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(TESTS, "golden")
GOLDEN_CASES = ("legacy39_5x4", "tiny4k_5x4", "mid50k_10x10", "small200k_8x8", "uniform1m_5x4")
# Thresholds of the golden vectors: 0 disables the exit; 75; 100 (cut 0);
# 99.99 (cut just above 0); 1 (cut 9,900); -5 (cut -2,100: exit at the first hit).
GOLDEN_THRESHOLDS = (0.0, 75.0, 100.0, 99.99, 1.0, -5.0)


def tkey(t: float) -> str:
    """Key of a threshold in the golden files (tests/golden/make_golden.py)."""
    return str(int(t)) if t == int(t) else repr(t)


def cut_of(t: float) -> float:
    """maximum_sum of lib/ISSLScoringModule.cpp:225."""
    return (10000.0 - t * 100) / t


def call_margin(o: dict, method: int, t: float):
    """Relative distance of the pass/fail test's operands from the threshold
    (lib/ISSLScoringModule.cpp:404-455), from oracle scores: a call can differ
    between sum orders only where this is within their rounding."""
    vm, vc = o["mit"], o["cfd"]
    if t == 0 or cut_of(t) == 0:
        # no call rejects at t = 0; at t = 100 a score is below 100 exactly
        # when its sum is positive, which no order of nonnegative terms changes
        return np.full(vm.shape, np.inf)
    if method in (2, 3):
        d = np.minimum(np.abs(vm - t), np.abs(vc - t))
    elif method == 4:
        d = np.abs((vm + vc) / 2 - t)
    elif method == 0:
        d = np.abs(vm - t)
    else:
        d = np.abs(vc - t)
    return d / abs(t)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


def load_golden(name):
    return np.load(os.path.join(GOLDEN, f"golden_{name}.npz"), allow_pickle=False)


def golden_index(name):
    """Host arrays of a golden case's index (committed file or regenerated)."""
    import cracklingplusplus_gpu_amd as issl
    from cracklingplusplus_gpu_amd import workloads
    if name == "legacy39_5x4":
        from golden.make_golden import decode_legacy
        return decode_legacy(os.path.join(GOLDEN, "ref_test_genome_offTargets_indexed.issl"))
    if name == "tiny4k_5x4":
        return issl.read_index(os.path.join(GOLDEN, "tiny4k_5x4.issl"))
    sigs, occ, masks, _, _ = workloads.make(name)
    return issl.build_lists(sigs, occ, masks)


@pytest.fixture(scope="session")
def gpu_available():
    import cracklingplusplus_gpu_amd as issl
    if issl.device_count() <= 0:
        pytest.fail("no HIP device: GPU tests need an MI355X")
    return True


@pytest.fixture(scope="session")
def c1_index(gpu_available):
    """The 100 M-site, 10-slice index (genome100m_10x10) resident on device 0,
    its host arrays, its 10,000 guides and their kinds; shared by the
    full-size modules."""
    import cracklingplusplus_gpu_amd as issl
    from cracklingplusplus_gpu_amd import workloads
    threads = max(1, min(16, len(os.sched_getaffinity(0))))
    print(f"[gpu-full] building genome100m_10x10 ({threads} threads)", flush=True)
    sigs, occ, masks, guides, kinds = workloads.make("genome100m_10x10", threads=threads)
    hx = issl.build_lists(sigs, occ, masks, threads=threads)
    del sigs, occ
    print("[gpu-full] uploading genome100m_10x10", flush=True)
    ix = issl.Index.from_arrays(hx)
    yield hx, ix, guides, kinds
    ix.close()
