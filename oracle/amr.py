"""ORACLE (test infrastructure only): the ZZ flux-recovery mesh-error estimate and the element marking, FP64.

Only tests, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may import this module; the
product path never does.

Restated from the reference:

* ``fem3D::calculateMeshErrors`` (src/OpenParEM3D/fem3D.cpp:1162-1380) with ``OPEM_L2ZZErrorEstimator``
  (OPEM_L2ZZErrorEstimator.cpp:18-164): for the real and the imaginary part of the H field separately, the
  element flux is the curl-curl integrator's flux (1/mu_r) curl H (fem3D.cpp:1167: CurlCurlIntegrator(Inv_mur));
  it is L2-projected onto ND_p (smooth_flux_fes, fem3D.cpp:1170-1171) by the mass solve P s = b,
  b_a = sum_K int_K (curl H / mu_r) . phi_a; each element's error is the L1 norm int_K |s - flux| dV
  (norm_p = 1, fem3D.cpp:1177; ComputeElementLpDistance), the two parts are combined as sqrt(e_re^2 + e_im^2)
  (:1221-1224) and the larger error of every driven port is kept (:1263-1267).  The integral uses the collapsed
  Gauss-Jacobi rule of ``nd.tet_rule(p + 2)`` (exact for degree 2p + 3); MFEM's own rule differs, so element errors
  agree with the reference to the quadrature error of a non-polynomial integrand, not to rounding.
* The marking (fem3D.cpp:1300-1325): refinementCount = int(NE * fraction) (at least 1) elements of largest error;
  the largest error is reported (maxMeshError).
"""
from __future__ import annotations

import numpy as np
import scipy.sparse.linalg as spla

from . import assemble as asm
from . import nd


def _element_fields(space, coefs, nquad, chunk=4096):
    """Yields (element slice, phi (ne,nq,nd,3), curl, |det| * w (ne,nq)) in element chunks."""
    basis = nd.tet_local_basis(space.p)
    L, w = nd.tet_rule(nquad or space.p + 2)
    v_all = space.sorted_coords()
    for e0 in range(0, v_all.shape[0], chunk):
        sl = slice(e0, min(v_all.shape[0], e0 + chunk))
        g, det = nd.tet_gradients(v_all[sl])
        phi, curl = nd.eval_basis(basis, L, g)
        yield sl, phi, curl, np.abs(det)[:, None] * w[None, :]


def smooth_flux(space, h, mur_elem, nquad=None):
    """ND_p L2 projection s of the element flux curl(h) / mu_r (complex DOF vector)."""
    _, Me = asm.volume_matrices(space)
    P = asm._scatter(space.elem_dofs, Me, space.ndof).tocsc()     # unit ND mass
    P.sum_duplicates()
    b = np.zeros(space.ndof, np.complex128)
    for sl, phi, curl, wq in _element_fields(space, None, nquad):
        he = h[space.elem_dofs[sl]]                                    # (ne,nd)
        flux = np.einsum("eqak,ea->eqk", curl, he) / mur_elem[sl][:, None, None]
        be = np.einsum("eq,eqak,eqk->ea", wq, phi, flux)
        np.add.at(b, space.elem_dofs[sl].ravel(), be.ravel())
    lu = spla.splu(P)
    return lu.solve(b.real) + 1j * lu.solve(b.imag)


def zz_errors(space, hs, mur_elem, nquad=None):
    """Element errors (ne,) of the H fields ``hs`` (list of complex DOF vectors, one per driven port): the
    largest over the ports of sqrt(e_re^2 + e_im^2)."""
    out = np.zeros(space.elem_dofs.shape[0])
    for h in hs:
        s = smooth_flux(space, h, mur_elem, nquad)
        for sl, phi, curl, wq in _element_fields(space, None, nquad):
            flux = np.einsum("eqak,ea->eqk", curl, h[space.elem_dofs[sl]]) / mur_elem[sl][:, None, None]
            sv = np.einsum("eqak,ea->eqk", phi, s[space.elem_dofs[sl]])
            d = sv - flux
            ere = np.sum(wq * np.linalg.norm(d.real, axis=2), axis=1)
            eim = np.sum(wq * np.linalg.norm(d.imag, axis=2), axis=1)
            out[sl] = np.maximum(out[sl], np.sqrt(ere * ere + eim * eim))
    return out


def mark(errors, fraction):
    """Indices of the int(NE * fraction) (at least 1) largest errors, largest first (ties: lower index first),
    and the largest error."""
    n = max(1, int(errors.size * fraction))
    order = np.lexsort((np.arange(errors.size), -errors))
    return order[:n], float(errors[order[0]])
