"""CPU checks of the conforming local refinement (opem_refine_mesh, host only; the step after the marking of
fem3D::calculateMeshErrors, which the reference hands to MFEM's GeneralRefinement).

Invariants that fix a correct refinement whatever the bisection order: every interior face is shared by exactly two
tets and the exterior faces are exactly the (split) boundary faces (conformity); volume per attribute and area
per boundary attribute are preserved; every tet keeps the orientation of its parent; no marked tet survives
unsplit; new vertices are edge midpoints of the input or of earlier bisections.
"""
import numpy as np
import pytest

from openparem_amd.amr import refine
from openparem_amd.mesh import microstrip_mesh, waveguide_mesh

FACES = ((0, 1, 2), (0, 1, 3), (0, 2, 3), (1, 2, 3))


def signed_vol(m):
    v = m.xyz[m.tets]
    return np.linalg.det(np.stack([v[:, 1] - v[:, 0], v[:, 2] - v[:, 0], v[:, 3] - v[:, 0]], 2)) / 6


def face_count(m):
    cnt = {}
    for t in np.sort(m.tets, axis=1):
        for f in FACES:
            k = tuple(t[list(f)])
            cnt[k] = cnt.get(k, 0) + 1
    return cnt


def tri_area(m, sel):
    v = m.xyz[m.bfaces[sel]]
    return 0.5 * np.linalg.norm(np.cross(v[:, 1] - v[:, 0], v[:, 2] - v[:, 0]), axis=1).sum()


@pytest.mark.parametrize("maker,frac", [(lambda: waveguide_mesh(nx=3, ny=2, nz=4, seed=2), 0.1),
                                        (lambda: waveguide_mesh(nx=3, ny=2, nz=4, seed=2), 1.0),
                                        (lambda: microstrip_mesh(nx=6, ny=4, nz=4, n_sub=2, seed=3), 0.25)])
def test_refinement_invariants(maker, frac):
    m = maker()
    rng = np.random.default_rng(4)
    marked = rng.permutation(m.nt)[:max(1, int(m.nt * frac))].astype(np.int32)
    r, par = refine(m, marked)
    assert r.nt > m.nt and r.nv > m.nv
    # conformity
    cnt = face_count(r)
    assert max(cnt.values()) <= 2
    ext = {k for k, c in cnt.items() if c == 1}
    bset = {tuple(sorted(f)) for f in r.bfaces.tolist()}
    ext0 = {k for k, c in face_count(m).items() if c == 1}
    bset0 = {tuple(sorted(f)) for f in m.bfaces.tolist()}
    if ext0 <= bset0:                                  # the input's exterior is all boundary faces
        assert ext <= bset
    assert all(k in cnt for k in bset)                 # every boundary face is a face of some tet
    # volume per attribute, area per boundary attribute, orientation
    v0, v1 = signed_vol(m), signed_vol(r)
    for a in np.unique(m.tet_attr):
        assert abs(np.abs(v1[r.tet_attr == a]).sum() - np.abs(v0[m.tet_attr == a]).sum()) < 1e-12 * np.abs(v0).sum()
    for a in np.unique(m.bface_attr):
        A0, A1 = tri_area(m, m.bface_attr == a), tri_area(r, r.bface_attr == a)
        assert abs(A1 - A0) < 1e-12 * A0
    assert np.all(np.sign(v1) == np.sign(v0[par])) and np.all(np.abs(v1) > 0)
    assert np.array_equal(r.tet_attr, m.tet_attr[par])
    # no marked tet survives
    surv = {tuple(sorted(t)) for t in r.tets.tolist()}
    assert not any(tuple(sorted(t)) in surv for t in m.tets[marked].tolist())
    # inputs vertices first and unchanged
    assert np.array_equal(r.xyz[:m.nv], m.xyz)


def test_bad_input_rejected():
    m = waveguide_mesh(nx=2, ny=1, nz=2, seed=1)
    with pytest.raises(ValueError):
        refine(m, np.array([m.nt], np.int32))
