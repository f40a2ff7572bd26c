import os
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
GOLDEN = Path(__file__).resolve().parent / "golden"
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


def load_golden(name: str):
    z = np.load(GOLDEN / name, allow_pickle=False)
    return {k: z[k] for k in z.files}


def golden_bfs_names():
    return sorted(p.name for p in GOLDEN.glob("bfs_*.npz"))


def symmetric_csr_from_edges(V, edges):
    adj = [set() for _ in range(V)]
    for u, v in edges:
        if u != v:
            adj[u].add(int(v))
            adj[v].add(int(u))
    beg = np.zeros(V + 1, dtype=np.int64)
    for v in range(V):
        beg[v + 1] = beg[v] + len(adj[v])
    a = np.array([w for v in range(V) for w in sorted(adj[v])], dtype=np.int32)
    return beg, a


@pytest.fixture(scope="session")
def gpu_available():
    import ibfs_amd
    return ibfs_amd.device_count() > 0
