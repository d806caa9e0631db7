"""ORACLE (test infrastructure only): one frequency point end to end, FP64.

Order of operations per (frequency, driving set) follows ``OpenParEM3D.cpp:497-555`` and
``fem3D.cpp:712-799``: build A, eliminate PEC, fill X for the driven port, solve E,
project H, restrict E/H to the ports, C+/C- splits; then S once per frequency
(``OpenParEM3D.cpp:558-566``).
"""
from __future__ import annotations

import numpy as np

from . import assemble as asm
from . import solve as slv


def te10_mode(mesh, attr, freq):
    """Analytic TE10 of a rectangular guide x in [0,a], y in [0,b] (waveguide_mesh geometry).

    E_t = y_hat sin(pi x/a);  H_t = (gamma/(j omega mu)) n_out x E_t, the wave leaving the domain
    through the port.  V = int_0^b E_y dy at x = a/2;  Z_PV = V V* / (2 conj(P)), P = 1/2 int E x H*.n.
    """
    a, b = mesh.geometry["a"], mesh.geometry["b"]
    er = mesh.materials[1]["er"] * (1 - 1j * mesh.materials[1]["tand"])
    mur = mesh.materials[1]["mur"]
    k0 = 2 * np.pi * freq * np.sqrt(asm.MU0 * asm.EPS0)
    kc = np.pi / a
    gamma = np.sqrt(kc * kc - k0 * k0 * er * mur + 0j)
    if gamma.imag < 0:
        gamma = -gamma
    if gamma.real < 0:
        gamma = -gamma
    omega = 2 * np.pi * freq
    n = mesh.port_normal[attr]

    def E(pts):
        e = np.zeros(pts.shape, np.complex128)
        e[..., 1] = np.sin(kc * pts[..., 0])
        return e

    def H(pts):
        return gamma / (1j * omega * asm.MU0 * mur) * np.cross(n, E(pts))

    V = b
    P = 0.5 * np.conj(gamma / (1j * omega * asm.MU0 * mur)) * (a * b / 2)   # 1/2 int (E x H*).n
    Z = 0.5 * V * np.conj(V) / np.conj(P)
    return dict(E=E, H=H, gamma=gamma, V=V, Z=Z)


class Problem:
    """Space + base data for one mesh/order; `solve_point` runs one frequency."""

    def __init__(self, mesh, p):
        self.mesh = mesh
        self.space = asm.Space(mesh, p)
        self.Ke, self.Me = asm.volume_matrices(self.space)
        self.ports = sorted(a for a, k in mesh.bdr_kind.items() if k == "port")
        pec_attrs = [a for a, k in mesh.bdr_kind.items() if k == "pec"]
        self.pec = self.space.dofs_on_bdr(pec_attrs)
        self.traces = {a: slv.PortTrace(self.space, a) for a in self.ports}

    def solve_point(self, freq, modes, solver="direct", restart=50, rtol=1e-10, rs=None):
        sp_ = self.space
        gam = {a: modes[a]["gamma"] for a in self.ports}
        A = asm.system_matrix(sp_, freq, gam, rs_of_attr=rs, Ke=self.Ke, Me=self.Me)
        P, Q = asm.h_matrices(sp_, freq, Me=self.Me)
        Em, Hm, xm = {}, {}, {}
        for a in self.ports:
            tr = self.traces[a]
            ev = modes[a]["E"](tr.pts)
            hv = modes[a]["H"](tr.pts)
            pec_on_port = np.intersect1d(tr.port_dofs, self.pec)
            xm[a] = tr.project(ev, fixed_zero=pec_on_port)
            hdofs = tr.project(hv)
            xe = np.zeros(sp_.ndof, np.complex128)
            xe[tr.port_dofs] = xm[a]
            xh = np.zeros(sp_.ndof, np.complex128)
            xh[tr.port_dofs] = hdofs
            Em[a], Hm[a] = tr.field(xe), tr.field(xh)
        N = len(self.ports)
        Cp = np.zeros((N, N), np.complex128)
        Cm = np.zeros((N, N), np.complex128)
        info = []
        for i, drv in enumerate(self.ports):
            tr = self.traces[drv]
            At, b = slv.excite(A, self.pec, tr.port_dofs, xm[drv])
            if solver == "direct":
                x = slv.direct_solve(At, b)
                its, ok = 0, True
            else:
                dmask = np.zeros(sp_.ndof, bool)
                dmask[self.pec] = True
                dmask[tr.port_dofs] = True
                if solver == "gmres-bj":
                    M = slv.BlockJacobi(At, sp_.elem_dofs, dmask)
                elif solver == "gmres-patch":
                    M = slv.VertexPatchASM(At, sp_, dmask)
                else:
                    M = None
                x, its, ok, _ = slv.gmres(At, b, M=M, restart=restart, rtol=rtol)
            h = slv.h_field(P, Q, x)
            res = np.linalg.norm(b - At @ x) / np.linalg.norm(b)
            info.append(dict(iters=its, converged=ok, true_rel_res=res, x=x, h=h))
            for k, a in enumerate(self.ports):
                trk = self.traces[a]
                Cp[i, k], Cm[i, k] = slv.splits(trk, Em[a], Hm[a], trk.field(x), trk.field(h))
        V = [modes[a]["V"] for a in self.ports]
        Z = [modes[a]["Z"] for a in self.ports]
        S = slv.calculate_s(Cp, Cm, V, Z)
        return dict(S=slv.force_reciprocal(S), S_raw=S, Cp=Cp, Cm=Cm, info=info)
