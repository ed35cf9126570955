"""Write size-independent summaries of the oracle's results on the full-size
R-MAT graphs (same seeded generator as the product), for GPU tests at sizes
where running the oracle inside a test would take minutes:

* rmat20_truss: kmax, histogram and a position-weighted checksum of the
  truss numbers (per entry of the symmetric pattern) from the oracle's
  independent peeling algorithm (oracle/oracle.cpp, or_truss_peel);
* rmat20_tc / rmat22_tc: triangle counts from the parallel MSA restatement.

    python tests/golden/make_large.py      # about 4 minutes on 8 cores
"""
import json
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent.parent))

from graphkokkos_amd import HostGraph  # noqa: E402
from oracle import oracle as O  # noqa: E402


def checksum(a: np.ndarray) -> int:
    w = (np.arange(a.size, dtype=np.uint64) % np.uint64(1000003)) + np.uint64(1)
    return int(np.sum(a.astype(np.uint64) * w, dtype=np.uint64))


def main():
    out = {}
    G = HostGraph.rmat(20, 16, seed=1)
    out["rmat20"] = {"edges": G.edges, "triangles": O.tc_sandia(G.L)}
    tr, kmax = O.truss_peel(G.A)
    out["rmat20"].update({"kmax": kmax, "truss_hist": np.bincount(tr).tolist(), "truss_checksum": checksum(tr)})
    print("rmat20", out["rmat20"]["edges"], out["rmat20"]["triangles"], kmax, flush=True)
    G = HostGraph.rmat(22, 16, seed=1)
    out["rmat22"] = {"edges": G.edges, "triangles": O.tc_sandia(G.L)}
    print("rmat22", out["rmat22"], flush=True)
    (HERE / "rmat_large.json").write_text(json.dumps(out))


if __name__ == "__main__":
    main()
