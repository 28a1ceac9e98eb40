import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "dpm-g4cpp_amd", "python"), os.path.join(ROOT, "dpm-g4cpp_amd", "tablegen"),
          os.path.join(ROOT, "tests"), ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP engine)")


@pytest.fixture(scope="session")
def tables3():
    import __graft_entry__
    return __graft_entry__.tables_dir(3)


@pytest.fixture(scope="session")
def tables0():
    import __graft_entry__
    return __graft_entry__.tables_dir(0)


@pytest.fixture(scope="session")
def tables1():
    import __graft_entry__
    return __graft_entry__.tables_dir(1)


@pytest.fixture(scope="session")
def tables2():
    import __graft_entry__
    return __graft_entry__.tables_dir(2)


@pytest.fixture(scope="session")
def gpu_tables3(tables3):
    import dpm_amd
    t = dpm_amd.Tables(tables3)
    yield t
    t.close()


@pytest.fixture(scope="session")
def gpu_tables0(tables0):
    import dpm_amd
    t = dpm_amd.Tables(tables0)
    yield t
    t.close()
