"""Shared fixtures.  GPU tests are marked `gpu`; everything else runs on CPU."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

ORACLE_LIB = os.path.join(ROOT, "oracle", "liboracle.so")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def tables():
    from gpet_amd import engine
    return engine.load_tables()


def oracle_engine(opt):
    from gpet_amd import engine
    return engine.Engine(opt, lib_path=ORACLE_LIB, prefix="gpet_cpu_")


def hip_engine(opt):
    from gpet_amd import engine
    return engine.Engine(opt)


def same_bytes(a, b):
    """Bitwise equality of two record arrays."""
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    return a.shape == b.shape and a.dtype == b.dtype and np.array_equal(a.view(np.uint8), b.view(np.uint8))


def first_diff(a, b):
    """Human-readable first differing record (for assertion messages)."""
    n = min(len(a), len(b))
    for i in range(n):
        if a[i].tobytes() != b[i].tobytes():
            return "len %d vs %d; first diff at %d:\n  gpu    %s\n  oracle %s" % (len(a), len(b), i, a[i], b[i])
    return "len %d vs %d" % (len(a), len(b))
