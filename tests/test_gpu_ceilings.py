"""GPU: the plain read / write / copy probe returns six positive rates and
leaves the device usable (a spindown context created afterwards works)."""
import math

import pytest

from tdfstat_amd import search as TS

pytestmark = pytest.mark.gpu


def test_memory_ceilings_are_positive_rates():
    print("memory ceilings 64 MiB x 2", flush=True)
    c = TS.memory_ceilings(64 << 20, 2)
    assert set(c["all"]) == {"read", "read_nt", "write", "write_nt", "copy", "copy_nt"}
    for k, v in c["all"].items():
        assert math.isfinite(v) and v > 10.0, (k, v)  # GB/s; far above any PCIe rate
    assert c["read"] == max(c["all"]["read"], c["all"]["read_nt"])
    assert c["buffer_bytes"] == 64 << 20
