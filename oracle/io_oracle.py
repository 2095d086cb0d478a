"""CPU ORACLE — test infrastructure only (tests/ use it as the checker).

A literal restatement of the reference's input readers:
parse_supported_msh2_mesh + build_structured_grid
(onelab_interface/gmsh_structured_grid.hpp:162-553: whitespace tokens, a
linear scan for the first axis value within 1e-12, node tags in a dict) and
the .silpro lexer / parser (onelab_interface/onelab_interface.hpp:167-389).
Errors raise ValueError with the reference's message.
"""
COORD_TOL = 1.0e-12

DIMENSION = {15: 0, 1: 1, 2: 2, 3: 2, 4: 3, 5: 3, 6: 3, 7: 3}
NODES = {1: 2, 2: 3, 3: 4, 4: 4, 5: 8, 6: 6, 7: 5, 15: 1}
TOPOLOGY = ("unsupported mesh topology: SimiLie currently requires the whole mesh to be made of "
            "quadrilaterals or hexahedra")


def _near(a, b):
    return abs(a - b) < COORD_TOL


def _unique_sorted(values):
    out = []
    for v in sorted(values):
        if not out or not _near(out[-1], v):
            out.append(v)
    return out


def _nearest(axis, v):
    for i, a in enumerate(axis):
        if _near(a, v):
            return i
    raise ValueError("coordinate does not belong to the detected structured grid")


def parse_msh2(text):
    """(nodes [(tag, x, y, z)], cells [(physical tag, node tags)], cell dim)"""
    toks = text.split()
    pos = 0

    def take():
        nonlocal pos
        if pos >= len(toks):
            raise ValueError("unexpected end of mesh file")
        pos += 1
        return toks[pos - 1]

    nodes, quads, hexes = [], [], []
    dim = 0
    saw = False
    while pos < len(toks):
        t = take()
        if t == "$MeshFormat":
            version, ftype = float(take()), int(take())
            take()
            if version < 2.0 or version >= 3.0 or ftype != 0:
                raise ValueError("only Gmsh msh2 ASCII meshes are supported")
        elif t == "$Nodes":
            for _ in range(int(take())):
                nodes.append((int(take()), float(take()), float(take()), float(take())))
        elif t == "$Elements":
            saw = True
            for _ in range(int(take())):
                take()
                etype, ntags = int(take()), int(take())
                tags = [int(take()) for _ in range(ntags)]
                if etype not in DIMENSION:
                    raise ValueError(f"unsupported Gmsh element type {etype} found in mesh")
                if etype in (3, 5):
                    d = 2 if etype == 3 else 3
                    if dim == 0:
                        dim = d
                    elif dim != d:
                        raise ValueError(TOPOLOGY)
                    cell = (tags[0] if tags else 0, [int(take()) for _ in range(NODES[etype])])
                    (quads if etype == 3 else hexes).append(cell)
                elif etype in (1, 2, 15):
                    for _ in range(NODES[etype]):
                        take()
                else:
                    for _ in range(NODES[etype]):
                        take()
                    raise ValueError(TOPOLOGY)
    if not nodes:
        raise ValueError("the provided mesh does not contain any node")
    if not saw or (not quads and not hexes):
        raise ValueError(TOPOLOGY)
    return nodes, (quads if quads else hexes), (2 if quads else 3)


def build_structured_grid(nodes, cells, dim):
    """(axis coordinates, cell tags x fastest, plane z)"""
    axes = [_unique_sorted([n[1 + a] for n in nodes]) for a in range(3)]
    plane = 0.0
    if dim == 2:
        if len(axes[2]) != 1:
            raise ValueError("the quadrilateral mesh nodes do not lie on a single plane")
        plane = axes[2][0]
        axes = axes[:2]
    shape = [len(a) for a in axes]
    total = 1
    for s in shape:
        total *= s
    ordered = [None] * total
    by_tag = {}
    for tag, x, y, z in nodes:
        idx = [_nearest(axes[a], (x, y, z)[a]) for a in range(dim)]
        lin = idx[0] + shape[0] * (idx[1] + (shape[1] * idx[2] if dim == 3 else 0))
        if ordered[lin] is not None:
            r = ordered[lin]
            if not (_near(r[1], x) and _near(r[2], y) and _near(r[3], z)):
                raise ValueError("duplicated nodes on the detected rectilinear grid")
            by_tag.setdefault(tag, idx)
            continue
        by_tag.setdefault(tag, idx)
        ordered[lin] = (tag, x, y, z)
    if any(o is None for o in ordered):
        raise ValueError("the mesh nodes do not form a full rectilinear grid")
    cshape = [s - 1 for s in shape]
    ncells = 1
    for s in cshape:
        ncells *= s
    kind, one = ("hexahedral", "a hexahedron") if dim == 3 else ("quadrilateral", "a quadrilateral")
    if len(cells) != ncells:
        raise ValueError(f"the {kind} cells do not cover a full rectilinear grid")
    tags = [None] * ncells
    for phys, ntags in cells:
        lo, hi = [10 ** 18] * dim, [0] * dim
        for t in ntags:
            if t not in by_tag:
                raise ValueError(f"{one} references an unknown node")
            for a in range(dim):
                lo[a] = min(lo[a], by_tag[t][a])
                hi[a] = max(hi[a], by_tag[t][a])
        if any(hi[a] != lo[a] + 1 for a in range(dim)):
            raise ValueError(f"{one} does not match a single rectilinear grid cell")
        lin = lo[0] + cshape[0] * (lo[1] + (cshape[1] * lo[2] if dim == 3 else 0))
        if tags[lin] is not None:
            raise ValueError(f"duplicated {'hexahedron' if dim == 3 else 'quadrilateral'} on the detected "
                             "rectilinear grid")
        tags[lin] = phys
    return axes, tags, plane


def read_structured_grid(text):
    return build_structured_grid(*parse_msh2(text))


# ------------------------------------------------------------------ silpro ---

def lex_silpro(s):
    toks = []
    i = 0
    while i < len(s):
        c = s[i]
        if c.isspace():
            i += 1
        elif c == "#" or (c == "/" and i + 1 < len(s) and s[i + 1] == "/"):
            while i < len(s) and s[i] != "\n":
                i += 1
        elif c in "{};":
            toks.append((c, c))
            i += 1
        elif c == '"':
            i += 1
            text = []
            while i < len(s) and s[i] != '"':
                if s[i] == "\\" and i + 1 < len(s):
                    i += 1
                text.append(s[i])
                i += 1
            if i == len(s):
                raise ValueError("unterminated string literal in .silpro file")
            i += 1
            toks.append(("string", "".join(text)))
        else:
            text = []
            while i < len(s) and not s[i].isspace() and s[i] not in "{};" and not (
                    s[i] == "/" and i + 1 < len(s) and s[i + 1] == "/"):
                text.append(s[i])
                i += 1
            if text:
                toks.append(("word", "".join(text)))
    return toks


def parse_silpro(text):
    """{"values": {...}, "sections": {...}} with the first of repeated keys kept"""
    toks = lex_silpro(text)
    pos = 0

    def peek():
        if pos >= len(toks):
            raise ValueError("unexpected end of .silpro file")
        return toks[pos]

    def name():
        nonlocal pos
        k, t = peek()
        if k not in ("word", "string"):
            raise ValueError("expected identifier in .silpro file")
        pos += 1
        return t

    def expect(kind, msg):
        nonlocal pos
        if peek()[0] != kind:
            raise ValueError(msg)
        pos += 1

    def body():
        expect("{", "expected '{' after section name in .silpro file")
        sec = {"values": {}, "sections": {}}
        while peek()[0] != "}":
            n = name()
            if peek()[0] == "{":
                sub = body()
                sec["sections"].setdefault(n, sub)
                continue
            v = name()
            expect(";", "expected ';' after assignment in .silpro file")
            sec["values"].setdefault(n, v)
        expect("}", "expected '}' in .silpro file")
        return sec

    root = {"values": {}, "sections": {}}
    while pos < len(toks):
        n = name()
        root["sections"].setdefault(n, body())
    return root
