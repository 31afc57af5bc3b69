import json
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

GOLDEN = json.loads((ROOT / "tests" / "golden" / "vectors.json").read_text())

NP = {
    "u32": np.uint32,
    "i32": np.int32,
    "f32": np.float32,
    "u64": np.uint64,
    "i64": np.int64,
    "f64": np.float64,
}


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


def golden(name: str) -> dict:
    for v in GOLDEN["vectors"]:
        if v["name"] == name:
            return v
    raise KeyError(name)


def golden_of(op: str) -> list[dict]:
    return [v for v in GOLDEN["vectors"] if v["op"] == op]


def keys_of(v: dict) -> np.ndarray:
    """Keys of a golden vector; float keys are stored as raw bits."""
    dt = np.dtype(NP[v["dtype"]])
    if "key_bits" in v:
        ut = np.uint32 if dt.itemsize == 4 else np.uint64
        return np.array(v["key_bits"], dtype=ut).view(dt)
    return np.array(v["keys"], dtype=dt)


@pytest.fixture(scope="session")
def exec_():
    """One executor on cuda:0 for the GPU tests (fails loudly without libmsv)."""
    import massively_amd as m

    ex = m.Executor(0)
    yield ex
    ex.sync()
    ex.close()
