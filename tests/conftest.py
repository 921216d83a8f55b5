import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
for p in (ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built HIP library")


@pytest.fixture(scope="session")
def orc():
    import oracle
    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def tiny(orc):
    return orc.configure(orc.parse_hmm_file(os.path.join(GOLDEN, "tiny.hmm"))[0])


@pytest.fixture(scope="session")
def pair_profiles(orc):
    return [orc.configure(h) for h in orc.parse_hmm_file(os.path.join(GOLDEN, "hmmer34_pair.hmm"))]


@pytest.fixture(scope="session")
def pair_targets(orc):
    names, seqs = [], []
    with open(os.path.join(GOLDEN, "hmmer34_targets.faa")) as f:
        for line in f:
            if line.startswith(">"):
                names.append(line[1:].strip())
                seqs.append(b"")
            elif names:
                seqs[-1] += line.strip().encode()
    return names, [orc.encode(s) for s in seqs]


def load_package():
    """Import the sentinel-x_amd package (its directory name is not a Python identifier)."""
    import importlib.util
    if "sentinel_x_amd" in sys.modules:
        return sys.modules["sentinel_x_amd"]
    pkg_dir = os.path.join(ROOT, "sentinel-x_amd")
    spec = importlib.util.spec_from_file_location("sentinel_x_amd", os.path.join(pkg_dir, "__init__.py"),
                                                  submodule_search_locations=[pkg_dir])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["sentinel_x_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


@pytest.fixture(scope="session")
def sxa():
    return load_package()


@pytest.fixture(scope="session")
def gctx(sxa):
    ctx = sxa.Context(0)
    yield ctx
    ctx.close()
