"""Ordering, symbolic LU and the compact slot plan.

Restates, on index sets only, what the reference does with SymPy matrices:

* ``SparsityOptimizer.compute_rcm_ordering`` (``src/mkpp/lowering.py:1145-1176``):
  networkx Cuthill-McKee on the symmetrised structure, reversed, with an
  identity fallback when the bandwidth would grow;
* ``detect_blocks`` / ``_check_block_independence`` (``:1178-1219``): Tarjan
  SCCs; a block-diagonal system is factorised per block with no permutation
  (``:734-772, 775-845``);
* ``_compute_lu_core`` (``:894-996``): row-wise Doolittle fill on
  W = I/(gamma h) - J;
* ``compact_plan`` (``src/mkpp/compact_plan.py:6-70``): J slots in physical
  row-major order, W slots = sorted J locations plus the diagonal, LU slots =
  sorted L/U locations, and for each pivot row the U row then the L column,
  each with its products in ascending inner index.

networkx is a generation-time dependency only: the resulting permutation is
written into the committed plan JSON.
"""

from __future__ import annotations

from dataclasses import dataclass, field


@dataclass
class Plan:
    n: int
    permutation: list[int]  # logical position -> physical species index
    j_entries: list[tuple[int, int]]  # physical (i, j) per J slot
    j_slots: dict  # logical (i, j) -> J slot
    w_locations: list[tuple[int, int]]  # logical, sorted
    w_slots: dict
    lu_locations: list[tuple[int, int]]  # logical, sorted
    lu_slots: dict
    operations: list[dict]  # {output, w, begin, end, diagonal}
    products: list[tuple[int, int]]
    lower: list[tuple[int, int]]
    upper: list[tuple[int, int]]
    lower_offsets: list[int]
    upper_offsets: list[int]
    diagonals: list[int]
    fill_in: int
    blocks: list[list[int]] = field(default_factory=list)
    block_diagonal: bool = False

    @property
    def inverse(self) -> list[int]:
        inv = [0] * self.n
        for logical, phys in enumerate(self.permutation):
            inv[phys] = logical
        return inv


def _structure_set(n: int, entries: list[tuple[int, int]]) -> set:
    # Same construction order as lowering.py:581-586 so that set iteration
    # (and therefore networkx adjacency order) matches.
    nz = set(entries)
    s = set()
    for i in range(n):
        for j in range(n):
            if (i, j) in nz:
                s.add((i, j))
    return s


def rcm_ordering(n: int, structure: set) -> list[int]:
    import networkx as nx

    G = nx.Graph()
    G.add_nodes_from(range(n))
    for i, j in structure:
        if i != j:
            G.add_edge(i, j)
    if G.number_of_edges() == 0:
        return list(range(n))
    cm = list(nx.utils.rcm.cuthill_mckee_ordering(G))
    rcm = list(reversed(cm))

    def bw(order):
        inv = [0] * n
        for new, old in enumerate(order):
            inv[old] = new
        return max(abs(inv[i] - inv[j]) for i, j in structure) if structure else 0

    if bw(rcm) > bw(list(range(n))):
        return list(range(n))
    return rcm


def predict_fill(n: int, structure: set) -> set:
    active = set(structure) | {(i, i) for i in range(n)}
    fill = set()
    for k in range(n):
        rows = [i for i in range(k + 1, n) if (i, k) in active]
        cols = [j for j in range(k + 1, n) if (k, j) in active]
        for i in rows:
            for j in cols:
                if (i, j) not in active:
                    fill.add((i, j))
                    active.add((i, j))
    return fill


def detect_blocks(n: int, structure: set) -> list[list[int]]:
    import networkx as nx

    G = nx.DiGraph()
    G.add_nodes_from(range(n))
    for i, j in structure:
        if i != j:
            G.add_edge(i, j)
    sccs = list(nx.strongly_connected_components(G))
    return sorted([sorted(s) for s in sccs], key=lambda b: b[0])


def blocks_independent(n: int, structure: set, blocks: list[list[int]]) -> bool:
    if len(blocks) <= 1:
        return False
    where = {}
    for b, blk in enumerate(blocks):
        for s in blk:
            where[s] = b
    allpos = structure | predict_fill(n, structure) | {(i, i) for i in range(n)}
    for i, j in allpos:
        if i != j and where.get(i) != where.get(j):
            return False
    return True


def doolittle_pattern(n: int, w_nz: set) -> tuple[set, set, int]:
    """L (strict lower) and U (upper incl. diagonal) locations with fill.

    Row i: U(i, j>=i) if W(i,j) or some k<i has L(i,k) and U(k,j); then
    L(j>i, i) likewise (lowering.py:925-968).
    """
    nz_L = {(i, i) for i in range(n)}  # the reference seeds the diagonal here
    nz_U = set()
    fill = 0
    lower = set()
    for i in range(n):
        for j in range(i, n):
            sub = any((i, k) in nz_L and (k, j) in nz_U for k in range(i))
            has_w = (i, j) in w_nz
            if has_w or sub:
                if not has_w:
                    fill += 1
                nz_U.add((i, j))
        if (i, i) not in nz_U:
            raise ValueError(f"structurally zero pivot at logical row {i}")
        for j in range(i + 1, n):
            sub = any((j, k) in nz_L and (k, i) in nz_U for k in range(i))
            has_w = (j, i) in w_nz
            if has_w or sub:
                if not has_w:
                    fill += 1
                nz_L.add((j, i))
                lower.add((j, i))
    return lower, nz_U, fill


def build_plan(n: int, j_entries: list[tuple[int, int]]) -> Plan:
    """Full reference-order plan from the physical J structure."""
    structure = _structure_set(n, j_entries)
    perm = rcm_ordering(n, structure)
    blocks = detect_blocks(n, structure)
    block_diag = blocks_independent(n, structure, blocks)

    if block_diag:
        # Per-block factorisation in physical order (lowering.py:775-845).
        order = list(range(n))
        L, U, fill = set(), set(), 0
        for blk in blocks:
            local_nz = {(a, b) for a, gi in enumerate(blk) for b, gj in enumerate(blk) if (gi, gj) in structure}
            local_nz |= {(a, a) for a in range(len(blk))}
            l, u, f = doolittle_pattern(len(blk), local_nz)
            L |= {(blk[a], blk[b]) for a, b in l}
            U |= {(blk[a], blk[b]) for a, b in u}
            fill += f
    else:
        order = perm
        inv = [0] * n
        for logical, phys in enumerate(order):
            inv[phys] = logical
        w_nz = {(inv[i], inv[j]) for i, j in structure} | {(i, i) for i in range(n)}
        L, U, fill = doolittle_pattern(n, w_nz)

    return compact(n, j_entries, order, L, U, fill, blocks, block_diag)


def compact(n, j_entries, order, L, U, fill, blocks=None, block_diag=False) -> Plan:
    inv = {phys: logical for logical, phys in enumerate(order)}
    j_slots = {(inv[i], inv[j]): s for s, (i, j) in enumerate(j_entries)}
    w_locations = sorted(set(j_slots) | {(i, i) for i in range(n)})
    w_slots = {loc: s for s, loc in enumerate(w_locations)}
    lu_locations = sorted(L | U)
    lu_slots = {loc: s for s, loc in enumerate(lu_locations)}
    operations, products = [], []
    for row in range(n):
        locs = [(row, c) for c in range(row, n) if (row, c) in U]
        locs += [(lo, row) for lo in range(row + 1, n) if (lo, row) in L]
        for i, j in locs:
            begin = len(products)
            for k in range(row):
                if (i, k) in L and (k, j) in U:
                    products.append((lu_slots[i, k], lu_slots[k, j]))
            operations.append(
                {
                    "output": lu_slots[i, j],
                    "w": w_slots.get((i, j), -1),
                    "begin": begin,
                    "end": len(products),
                    "diagonal": lu_slots[row, row] if i != row else -1,
                }
            )
    lower, upper, lo_off, up_off = [], [], [0], [0]
    for row in range(n):
        lower.extend((c, lu_slots[row, c]) for c in range(row) if (row, c) in L)
        upper.extend((c, lu_slots[row, c]) for c in range(row + 1, n) if (row, c) in U)
        lo_off.append(len(lower))
        up_off.append(len(upper))
    return Plan(
        n=n,
        permutation=list(order),
        j_entries=list(j_entries),
        j_slots=j_slots,
        w_locations=w_locations,
        w_slots=w_slots,
        lu_locations=lu_locations,
        lu_slots=lu_slots,
        operations=operations,
        products=products,
        lower=lower,
        upper=upper,
        lower_offsets=lo_off,
        upper_offsets=up_off,
        diagonals=[lu_slots[i, i] for i in range(n)],
        fill_in=fill,
        blocks=blocks or [],
        block_diagonal=block_diag,
    )
