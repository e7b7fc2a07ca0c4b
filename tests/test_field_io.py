"""exnode/exelem export (FieldExport.c layout) and the ndiff comparison."""
import os

import numpy as np
import pytest

from cm_amd import field_io as fio
from cm_amd.mesh import RegularMesh, as_user_mesh


class _Set:
    """Stands in for an equations set: a mesh and its DOF vector."""

    def __init__(self, mesh, dofs):
        self.mesh, self._d = mesh, dofs

    def get_dofs(self):
        return self._d


def _mr_set(mesh, seed=0):
    rng = np.random.default_rng(seed)
    d = np.concatenate([mesh.coords.ravel() + 0.01 * rng.standard_normal(3 * mesh.n_nodes),
                        rng.standard_normal(mesh.n_nodes_l)])
    return _Set(mesh, d)


def test_exnode_round_trip_and_padding(tmp_path):
    mesh = RegularMesh.generate(2, (2, 1, 1), jitter=0.1, seed=1)
    s = _mr_set(mesh)
    base = os.path.join(tmp_path, "mr")
    fio.export_mooney_rivlin(base, s)
    nodes, headers = fio.read_exnode(base + ".part0.exnode")
    assert sorted(nodes) == list(range(1, mesh.n_nodes + 1))
    nq = mesh.n_nodes
    d = s.get_dofs()
    # corner (i, j, k) of the pressure component is quadratic grid node (2i, 2j, 2k)
    corner = {}
    n1, n2, _ = mesh.n
    T1, T2 = 2 * n1 + 1, 2 * n2 + 1
    for m in range(mesh.n_nodes_l):
        ci, cj, ck = m % (n1 + 1), (m // (n1 + 1)) % (n2 + 1), m // ((n1 + 1) * (n2 + 1))
        corner[2 * ci + 2 * cj * T1 + 2 * ck * T1 * T2] = m
    for p in range(nq):
        v = nodes[p + 1]
        assert len(v) == 7
        assert v[:3] == list(mesh.coords[:, p])                 # %.16E round-trips exactly
        assert v[3:6] == [d[p], d[nq + p], d[2 * nq + p]]
        assert v[6] == (d[3 * nq + corner[p]] if p in corner else fio.PADDING)
    h = headers[0]
    assert h[0] == "#Fields=2" and h[1] == "1) Geometry, coordinate, rectangular cartesian, #Components=3"
    assert h[5] == "2) Dependent, field,  rectangular cartesian, #Components=4"
    assert h[9] == "4.  Value index= 7, #Derivatives= 0"


def test_exnode_user_numbers_and_laplace_layout(tmp_path):
    m = RegularMesh.generate(1, (2, 2, 1))
    rng = np.random.default_rng(4)
    u = as_user_mesh(m, region_order=rng.permutation(m.n_nodes), user_numbers=np.arange(m.n_nodes) * 7 + 100)
    phi = rng.standard_normal(u.n_nodes)
    base = os.path.join(tmp_path, "lap")
    fio.export_laplace(base, _Set(u, phi))
    nodes, headers = fio.read_exnode(base + ".part0.exnode")
    assert len(headers) == 1 and headers[0][0] == "#Fields=2"
    for p, user in enumerate(u.node_user_numbers):
        assert nodes[int(user)] == list(u.coords[:, p]) + [phi[p]]
    text = open(base + ".part0.exelem").read().splitlines()
    assert text[:5] == [" Group name: LaplaceEquation", " Shape.  Dimension=3", " #Scale factor sets= 1",
                        " l.Lagrange*l.Lagrange*l.Lagrange, #Scale factors=8", " #Nodes=           8"]
    e0 = text.index(" Element:            1 0 0")
    assert [int(x) for x in text[e0 + 2].split()] == [int(x) for x in u.node_user_numbers[u.elem_nodes[0]]]


def test_exelem_mixed_bases(tmp_path):
    mesh = RegularMesh.generate(2, (1, 1, 1))
    base = os.path.join(tmp_path, "mr")
    fio.export_mooney_rivlin(base, _mr_set(mesh))
    text = open(base + ".part0.exelem").read().splitlines()
    assert text[2:5] == [" #Scale factor sets= 2", " q.Lagrange*q.Lagrange*q.Lagrange, #Scale factors=27",
                         " l.Lagrange*l.Lagrange*l.Lagrange, #Scale factors=8"]
    # the pressure component's eight nodes are the corners of the 27-node list
    k = text.index("   4.   l.Lagrange*l.Lagrange*l.Lagrange, no modify, standard node based.")
    idx = [int(text[k + 2 + 3 * j].split(".")[0]) for j in range(8)]
    assert idx == [1, 3, 7, 9, 19, 21, 25, 27]
    assert text[k + 4] == "      Scale factor indices:   28"


@pytest.mark.parametrize("a,b,tol,same", [
    ("  1.0000000000000000E+00\n", "  1.0000000000000000E+00\n", "0", True),
    ("  1.0000000000000000E+00\n", "  1.0000000000100000E+00\n", "0", False),
    ("  1.0000000000000000E+00\n", "  1.0000000000100000E+00\n", "1e-10", True),
    ("  1.0000000000000000E+00\n", "  1.0000000010000000E+00\n", "1e-10", False),
    # "%": relative to |a| when above 1 (1e-8% = 1e-10 of 2500 = 2.5e-7)
    ("  2.5000000000000000E+03\n", "  2.5000000001000000E+03\n", "1e-8%", True),
    ("  2.5000000000000000E+03\n", "  2.5000001000000000E+03\n", "1e-8%", False),
    # a change in the last printed digit is accepted once a tolerance is given
    (" x = 1.25 \n", " x = 1.26 \n", "1e-12", True),
    (" x = 1.25 \n", " x = 1.28 \n", "1e-12", False),
    (" Node:            3\n", " Node:            4\n", "1", False),
])
def test_ndiff_rule(tmp_path, a, b, tol, same):
    pa, pb = os.path.join(tmp_path, "a"), os.path.join(tmp_path, "b")
    open(pa, "w").write(a)
    open(pb, "w").write(b)
    assert (fio.ndiff(pa, pb, tol) == []) == same


def test_ndiff_of_two_exports(tmp_path):
    mesh = RegularMesh.generate(2, (1, 1, 2), jitter=0.1, seed=3)
    s = _mr_set(mesh, 5)
    fio.export_mooney_rivlin(os.path.join(tmp_path, "a"), s)
    d = s.get_dofs().copy()
    d[7] += 1e-13
    fio.export_mooney_rivlin(os.path.join(tmp_path, "b"), _Set(mesh, d))
    a, b = os.path.join(tmp_path, "a.part0.exnode"), os.path.join(tmp_path, "b.part0.exnode")
    assert len(fio.ndiff(a, b, "0")) == 1
    assert fio.ndiff(a, b, "1e-12") == []
