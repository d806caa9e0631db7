"""CPU checks of the named workloads the benchmark and the measurement driver run (no GPU)."""
import numpy as np

from openparem_amd import workloads
from openparem_amd.mesh import aspect_ratio


def test_named_workloads():
    w = workloads.make("waveguide-sweep")
    assert w["order"] == 2 and len(w["freqs"]) == 64
    assert w["freqs"][0] == 7e9 and w["freqs"][-1] == 12.5e9
    c4 = workloads.make("microstrip-c4")
    assert len(c4["freqs"]) == 64 and c4["freqs"][0] == 1e9 and c4["freqs"][-1] == 20e9
    assert 4 in c4["boundaries"]  # copper trace as a surface impedance


def test_c5_set():
    """Sixteen seeded instances: alternating guide/microstrip families, ND1/ND2 and one ND3, then 3-port tees and
    4-port crosses; 8-32 frequencies in each family's band; valid tetrahedra (positive volume, bounded aspect
    ratio)."""
    inst = workloads.c5_instances()
    assert len(inst) == 16
    assert [w["family"] for w in inst] == ["waveguide", "microstrip"] * 6 + ["tee", "cross"] * 2
    assert sorted({w["order"] for w in inst}) == [1, 2, 3] and sum(w["order"] == 3 for w in inst) == 1
    assert [len([k for k, v in w["mesh"].bdr_kind.items() if v == "port"]) for w in inst[12:]] == [3, 4, 3, 4]
    for w in inst:
        f = w["freqs"]
        assert len(f) in (8, 16, 24, 32)
        lo, hi = (1e9, 20e9) if w["family"] == "microstrip" else (7e9, 12.5e9)
        assert np.isclose(f[0], lo) and np.isclose(f[-1], hi)
        m = w["mesh"]
        a, b, c, d = (m.xyz[m.tets[:, k]] for k in range(4))
        vol = np.einsum("ij,ij->i", np.cross(b - a, c - a), d - a)
        assert (np.abs(vol) > 0).all()
        assert aspect_ratio(m.xyz[m.tets]).max() < 8.0
