"""GTH pseudopotential files in the reference's own text format.

``read_ps_gth`` follows ``read_ps_gth`` (``ps_gth_module.f90:16-160``), the
reader the reference selects for ``ippform`` 4 (``pseudopot_module.f90:
174-194``):

* line 1: ``name  Znuc  Zps`` (element name, nuclear charge, valence charge);
* line 2: ``rloc  C1  C2  C3  C4`` (local part);
* then one line per angular momentum l = 0, 1, ... (at most six):
  ``r_l  h_1^l  h_2^l  h_3^l``; every nonzero ``h_n^l`` is one projector
  (l, n) with radius r_l.

The reference counts the nonblank lines after line 2 (up to six): at most
two means the diagonal GTH form (``ps_type`` 0), more means the HGH form,
whose l >= 1 lines are followed by a line of spin-orbit ``k`` values and
whose projectors couple through off-diagonal ``h_nm`` (``calc_nondiagonal``,
``ps_type`` 1).  A first line containing ``KRACK`` selects Krack's HGH
layout.  This library implements the diagonal GTH operator only, so both
HGH layouts are rejected with an error instead of being applied as GTH.

Values are read list-directed: whitespace- or comma-separated numbers, and
Fortran ``d`` exponents are accepted.  A negative ``h`` keeps its sign here
(``GTHSpecies``); the reference stores ``|h|`` with ``inorm = -1``
(``ps_gth_module.f90:129-138``), which is the same projector sign.
"""
from __future__ import annotations

import os
import re

from .systems import GTHSpecies

LREFMAX = 6  # lines of projectors the reference looks at (lrefmax)


class UnsupportedPseudopotential(ValueError):
    """A file the reference reads as HGH (ps_type 1) or that is not GTH."""


def _numbers(line: str) -> list:
    toks = [t for t in re.split(r"[,\s]+", line.strip()) if t]
    out = []
    for t in toks:
        try:
            out.append(float(t.replace("d", "e").replace("D", "e")))
        except ValueError:
            break  # list-directed input ends at the first non-number (a comment)
    return out


def parse_gth(text: str, source: str = "<text>") -> GTHSpecies:
    """Parse the text of a GTH file (see the module docstring)."""
    lines = text.splitlines()
    if not lines:
        raise ValueError(f"{source}: empty pseudopotential file")
    if "KRACK" in lines[0]:
        raise UnsupportedPseudopotential(f"{source}: Krack-format HGH pseudopotential (ps_type 1) is not supported")
    if len(lines) < 2:
        raise ValueError(f"{source}: needs the 'name Znuc Zps' and 'rloc C1..C4' lines")
    head = lines[0].split()
    if len(head) < 3:
        raise ValueError(f"{source}: first line must be 'name Znuc Zps'")
    name = head[0].strip("'\"")
    try:
        znuc, zps = (float(v.replace("d", "e").replace("D", "e")) for v in head[1:3])
    except ValueError as e:
        raise ValueError(f"{source}: bad Znuc/Zps on line 1") from e
    loc = _numbers(lines[1])
    if len(loc) < 5:
        raise ValueError(f"{source}: line 2 must hold rloc and C1..C4")
    rloc, cloc = loc[0], tuple(loc[1:5])
    rest = lines[2:2 + LREFMAX]
    nonblank = sum(1 for ln in rest if ln.strip())
    if nonblank > 2:
        raise UnsupportedPseudopotential(
            f"{source}: {nonblank} projector lines: HGH format (ps_type 1, non-diagonal h_nm) is not supported")
    projectors = []
    # list-directed reads skip blank records, so the l-th nonblank line
    # after line 2 is the line of angular momentum l
    records = [ln for ln in lines[2:] if ln.strip()][:LREFMAX]
    for l, ln in enumerate(records):
        v = _numbers(ln)
        if len(v) < 4:
            raise ValueError(f"{source}: projector line for l = {l} needs r_l and h_1..h_3")
        r_l, h = v[0], v[1:4]
        for n, hn in enumerate(h, start=1):
            if hn != 0.0:
                projectors.append((l, n, r_l, hn))
    return GTHSpecies(name=name, zps=zps, rloc=rloc, cloc=cloc, projectors=tuple(projectors), znuc=znuc)


def read_ps_gth(path: str | os.PathLike) -> GTHSpecies:
    """Read a GTH pseudopotential file (``read_ps_gth``,
    ``ps_gth_module.f90:16-160``); raises UnsupportedPseudopotential for the
    HGH forms."""
    with open(path) as f:
        return parse_gth(f.read(), source=os.fspath(path))


def format_gth(sp: GTHSpecies) -> str:
    """Text of ``sp`` in the format read_ps_gth reads (at most one radius per
    l, as the format holds)."""
    lines = [f"{sp.name} {sp.znuc:.1f} {sp.zps:.1f}",
             " ".join(f"{v:.8f}" for v in (sp.rloc, *sp.cloc))]
    lmax = max((l for l, *_ in sp.projectors), default=-1)
    for l in range(lmax + 1):
        ps = [p for p in sp.projectors if p[0] == l]
        radii = {p[2] for p in ps}
        if len(radii) > 1:
            raise ValueError(f"l = {l} has projectors of different radii")
        h = [0.0, 0.0, 0.0]
        for _, n, _, hn in ps:
            h[n - 1] = hn
        r = radii.pop() if radii else 0.0
        lines.append(" ".join(f"{v:.8f}" for v in (r, *h)))
    return "\n".join(lines) + "\n"
