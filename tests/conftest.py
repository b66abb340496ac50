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
