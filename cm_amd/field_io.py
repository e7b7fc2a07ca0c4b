"""exnode / exelem export of nodal fields and an ndiff-style comparison.

The reference writes its results as CMISS exnode/exelem text (FieldExport.c,
driven by FIELD_IO_NODES_EXPORT / FIELD_IO_ELEMENTS_EXPORT,
field_IO_routines.f90:6158-6195, 5355-5595, 3533-3853) and compares
regression results with ``utils/ndiff``.  This module writes the same layout
for the fields of this package's equations sets:

exnode (FIELD_IO_EXPORT_NODES_INTO_LOCAL_FILE)
  `` Group name: <region>``, then for every run of nodes with the same
  component layout a header (FIELD_IO_EXPORT_NODAL_GROUP_HEADER_FORTRAN):
  `` #Fields=<variables>``, per variable `` i) <name>, <type label>,
  #Components=<n>`` and per component ``   <label>.  Value index= <k>,
  #Derivatives= 0`` (Lagrange bases carry no derivatives, one version); a
  component the node does not carry is padded with one value, 1.23456789
  (field_IO_routines.f90:5378).  Each node: `` Node:            <user>``
  and one line per component ``  %.16E`` (FieldExport_File_NodeValues).

exelem (FIELD_IO_EXPORT_ELEMENTS_INTO_LOCAL_FILE)
  `` Group name: <region>``, `` Shape.  Dimension=3``, then one header for
  the elements (they share one layout): scale factor sets (one per basis, one
  unit scale factor per element node), `` #Nodes=``, the fields with their
  basis per component and per-node value and scale-factor indices; each
  element: `` Element:            <user> 0 0``, its node user numbers and
  scale factors.

The type labels are FieldExport_GetVariableLabel / _GetCoordinateVariableLabel
(FieldExport.c:490-605): a geometric U variable is ``coordinate, rectangular
cartesian`` with components x, y, z; a general or geometric-general field's
variables are ``field,  rectangular cartesian``.
"""
import re

import numpy as np

PADDING = 1.23456789  # field_IO_routines.f90:5378

COORDINATE_LABEL = "coordinate, rectangular cartesian"
FIELD_LABEL = "field,  rectangular cartesian"


class Variable:
    """One exported field variable: ``values`` is (n_components, n_nodes) in the
    mesh's local node order; a component absent at a node holds NaN (the pressure
    of a mid-side node)."""

    def __init__(self, name, values, coordinate=False, label=None):
        self.name = name
        self.values = np.atleast_2d(np.asarray(values, np.float64))
        self.coordinate = coordinate
        self.label = label or (COORDINATE_LABEL if coordinate else FIELD_LABEL)

    def component_labels(self):
        n = self.values.shape[0]
        if self.coordinate and n <= 3:
            return ["x", "y", "z"][:n]
        return [str(i + 1) for i in range(n)]


def _fmt(v):
    return f"{v:.16E}"


def write_exnode(path, group, node_user_numbers, variables):
    """Writes the exnode file for nodes in local order (user numbers given)."""
    users = np.asarray(node_user_numbers)
    lines = [f" Group name: {group}"]
    prev_layout = None
    for p, user in enumerate(users):
        layout = tuple(tuple(bool(np.isfinite(v.values[c, p])) for c in range(v.values.shape[0])) for v in variables)
        if layout != prev_layout:
            lines.append(f" #Fields={len(variables)}")
            value_idx = 1
            for i, (v, present) in enumerate(zip(variables, layout)):
                lines.append(f" {i + 1}) {v.name}, {v.label}, #Components={v.values.shape[0]}")
                for label in v.component_labels():
                    lines.append(f"   {label}.  Value index= {value_idx}, #Derivatives= 0")
                    value_idx += 1
            prev_layout = layout
        lines.append(f" Node:            {int(user)}")
        for v, present in zip(variables, layout):
            for c, here in enumerate(present):
                lines.append("  " + _fmt(v.values[c, p] if here else PADDING))
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


_BASIS_LABEL = {1: "l.Lagrange*l.Lagrange*l.Lagrange", 2: "q.Lagrange*q.Lagrange*q.Lagrange"}


def write_exelem(path, group, element_user_numbers, bases, variables):
    """Writes the exelem file of hexahedral Lagrange elements.

    ``bases`` maps a basis order (1, 2) to the element node user numbers
    (E, (order+1)^3); ``variables`` is a list of (name, label, coordinate,
    component_orders): the basis order of every component."""
    orders = []  # scale factor sets in order of first use by a component
    for _, _, _, comp_orders in variables:
        for o in comp_orders:
            if o not in orders:
                orders.append(o)
    users = np.asarray(element_user_numbers)
    lines = [f" Group name: {group}", " Shape.  Dimension=3",
             f" #Scale factor sets= {len(orders)}"]
    for o in orders:
        lines.append(f" {_BASIS_LABEL[o]}, #Scale factors={(o + 1) ** 3}")
    max_nodes = max((o + 1) ** 3 for o in orders)
    lines.append(f" #Nodes=           {max_nodes}")
    lines.append(f" #Fields={len(variables)}")
    scale_start = {}
    s = 0
    for o in orders:
        scale_start[o] = s
        s += (o + 1) ** 3
    # the element node list is the basis with most nodes; a lower-order
    # component's nodes are located in it by user number
    main = max(orders)
    main_nodes = np.asarray(bases[main])
    for i, (name, label, coordinate, comp_orders) in enumerate(variables):
        lines.append(f" {i + 1}) {name}, {label}, #Components={len(comp_orders)}")
        labels = ["x", "y", "z"] if coordinate and len(comp_orders) <= 3 else [str(c + 1) for c in range(len(comp_orders))]
        for c, o in enumerate(comp_orders):
            lines.append(f"   {labels[c]}.   {_BASIS_LABEL[o]}, no modify, standard node based.")
            nn = (o + 1) ** 3
            lines.append(f"     #Nodes= {nn}")
            if o == main:
                node_index = np.arange(1, nn + 1)
            else:  # position of each of this basis' nodes in the main node list (first element)
                first = list(main_nodes[0])
                node_index = np.array([first.index(u) + 1 for u in np.asarray(bases[o])[0]])
            for k in range(nn):
                lines.append(f" {node_index[k]:5d}.  #Values=1")
                lines.append("      Value indices:     1")
                lines.append(f"      Scale factor indices: {scale_start[o] + k + 1:4d}")
    for e, user in enumerate(users):
        lines.append(f" Element:            {int(user)} 0 0")
        lines.append("   Nodes:")
        lines.append("".join(f"  {int(u):10d}" for u in main_nodes[e]))
        lines.append("   Scale factors:")
        for o in orders:  # one line per scale factor set, unit factors (Lagrange)
            lines.append("  " + "".join(f"   {_fmt(1.0)}" for _ in range((o + 1) ** 3)))
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def read_exnode(path):
    """Parses an exnode file written by write_exnode: returns
    {node user number: [values in value-index order]} and the headers seen."""
    nodes, headers, cur = {}, [], None
    with open(path) as f:
        for line in f:
            s = line.strip()
            if s.startswith("#Fields="):
                headers.append([s])
            elif headers and re.match(r"^\d+\)", s):
                headers[-1].append(s)
            elif headers and "Value index=" in s:
                headers[-1].append(s)
            elif s.startswith("Node:"):
                cur = int(s.split()[1])
                nodes[cur] = []
            elif cur is not None and s:
                nodes[cur].extend(float(x) for x in s.split())
    return nodes, headers


# --- ndiff (utils/ndiff) --------------------------------------------------------
_REAL = r"[+-]?(?:\d+\.|\.\d)\d*(?:[EeDd][+-]?\d+)?"
_SPLIT = re.compile(r"(^|[\]\[)(}{,;:'\"$\t ]) *(" + _REAL + r") *((?:[a-z][-a-z/^0-9]*|%)?\.?[\]\[)(}{,;:*'\"\t \n])")


def _real_close(a, b, tol, relative):
    """real_string_close: |b - a| <= tolerance, relative to |a| when above 1,
    and never finer than the resolution of a's printed digits."""
    a_s, b_s = a.replace("d", "e").replace("D", "e"), b.replace("d", "e").replace("D", "e")
    allowed = tol
    if relative:
        ref = abs(float(a_s))
        if ref > 1:
            allowed *= ref
    if tol > 0:
        m = re.search(r"\.(\d*)(?:[EeDd]([+-]?\d+))?", a_s)
        if not (m.group(2) is not None and float(a_s) == 0):
            exponent = -len(m.group(1)) + (int(m.group(2)) if m.group(2) is not None else 0)
            allowed = max(allowed, 1.1 * 10.0 ** exponent)
    return abs(float(b_s) - float(a_s)) <= allowed


def lines_match(a, b, tol=0.0, relative=False):
    """ncompare: the texts match and corresponding real numbers agree within
    the tolerance."""
    sa, sb = a, b
    while sa != sb:
        ma, mb = _SPLIT.search(sa), _SPLIT.search(sb)
        pa = sa[:ma.start()] + ma.group(1) if ma else sa
        pb = sb[:mb.start()] + mb.group(1) if mb else sb
        if pa != pb:
            return False
        if ma and mb:
            if not _real_close(ma.group(2), mb.group(2), tol, relative):
                return False
            sa, sb = ma.group(3) + sa[ma.end():], mb.group(3) + sb[mb.end():]
        elif ma or mb:
            return False
        else:
            return True
    return True


def ndiff(path_a, path_b, tolerance="0"):
    """Lines (number, line a, line b) that differ beyond the tolerance
    ("1e-10" absolute, "0.01%" relative).  Files of different lengths differ
    at the first missing line."""
    tol_s = str(tolerance)
    relative = tol_s.endswith("%")
    tol = float(tol_s[:-1]) / 100 if relative else float(tol_s)
    with open(path_a) as fa, open(path_b) as fb:
        la, lb = fa.readlines(), fb.readlines()
    out = []
    for i in range(max(len(la), len(lb))):
        a = la[i] if i < len(la) else ""
        b = lb[i] if i < len(lb) else ""
        if not lines_match(a, b, tol, relative):
            out.append((i + 1, a.rstrip("\n"), b.rstrip("\n")))
    return out


# --- the equations sets' fields ---------------------------------------------------
def export_laplace(basename, eq, group="LaplaceEquation"):
    """<basename>.part0.exnode / .exelem of a Laplace set: the geometry and the
    dependent field's U variable (its current DOFs)."""
    mesh = eq.mesh
    users = getattr(mesh, "node_user_numbers", None)
    users = np.arange(1, mesh.n_nodes + 1) if users is None else users
    u = eq.get_dofs()
    write_exnode(f"{basename}.part0.exnode", group, users,
                 [Variable("Geometry", mesh.coords, coordinate=True), Variable("Phi", u[None, :])])
    e_users = getattr(mesh, "element_user_numbers", None)
    e_users = np.arange(1, mesh.n_elements + 1) if e_users is None else e_users
    write_exelem(f"{basename}.part0.exelem", group, e_users, {1: users[mesh.elem_nodes]},
                 [("Geometry", COORDINATE_LABEL, True, [1, 1, 1]), ("Phi", FIELD_LABEL, False, [1])])


def export_mooney_rivlin(basename, eq, group="Region"):
    """<basename>.part0.exnode / .exelem of a Mooney-Rivlin set: the geometry
    and the dependent field (deformed coordinates on the triquadratic nodes,
    hydrostatic pressure on the trilinear corner nodes)."""
    mesh = eq.mesh
    nq, nl = mesh.n_nodes, mesh.n_nodes_l
    users = getattr(mesh, "node_user_numbers", None)
    if users is None:
        users = np.arange(1, nq + 1)
        corner_users = _regular_corner_users(mesh, users)
    else:
        corner_users = mesh.node_user_numbers_l
    d = eq.get_dofs()
    dep = np.full((4, nq), np.nan)
    dep[:3] = d[:3 * nq].reshape(3, nq)
    pos = {int(v): i for i, v in enumerate(users)}
    for m in range(nl):
        dep[3, pos[int(corner_users[m])]] = d[3 * nq + m]
    write_exnode(f"{basename}.part0.exnode", group, users,
                 [Variable("Geometry", mesh.coords, coordinate=True), Variable("Dependent", dep)])
    e_users = getattr(mesh, "element_user_numbers", None)
    e_users = np.arange(1, mesh.n_elements + 1) if e_users is None else e_users
    write_exelem(f"{basename}.part0.exelem", group, e_users,
                 {2: users[mesh.elem_nodes], 1: corner_users[mesh.elem_nodes_l]},
                 [("Geometry", COORDINATE_LABEL, True, [2, 2, 2]), ("Dependent", FIELD_LABEL, False, [2, 2, 2, 1])])


def _regular_corner_users(mesh, users):
    n1, n2, n3 = mesh.n
    T1, T2 = 2 * n1 + 1, 2 * n2 + 1
    i = np.arange(mesh.n_nodes_l)
    ci, cj, ck = i % (n1 + 1), (i // (n1 + 1)) % (n2 + 1), i // ((n1 + 1) * (n2 + 1))
    return users[2 * ci + 2 * cj * T1 + 2 * ck * T1 * T2]
