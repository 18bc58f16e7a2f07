import gzip
import os
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
for p in (ROOT, HERE):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(HERE, "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X); runs the native kernels")
    config.addinivalue_line("markers", "slow: long-running (full benchmark-size graphs)")


@pytest.fixture(scope="session")
def wiki_vote_path(tmp_path_factory):
    """The reference's dataset/wiki-vote_input (kept gzipped in tests/golden)."""
    d = tmp_path_factory.mktemp("wiki")
    p = d / "wiki-vote_input"
    with gzip.open(os.path.join(GOLDEN, "wiki-vote_input.gz"), "rb") as f:
        p.write_bytes(f.read())
    return str(p)


@pytest.fixture(scope="session")
def rmat13_path(tmp_path_factory):
    """R-MAT scale 13, edge factor 12, seed 1 written as a GraphPi text file
    (the file the reference counted for the golden values)."""
    import graphpi_amd as gp
    d = tmp_path_factory.mktemp("rmat13")
    p = d / "rmat13.txt"
    h = gp.HostGraph.rmat(13, 12, 0.57, 0.19, 0.19, 1)
    h.write_text(str(p))
    h.free()
    return str(p)


def load_json(name):
    import json
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)
