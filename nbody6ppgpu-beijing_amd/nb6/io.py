"""Input files of Nbody6++GPU: the NAMELIST parameter file (input.F:18-60, read with
Fortran namelist input) and the initial conditions on unit 10 (data.F:36-66).

* ``read_namelist`` / ``write_namelist``: groups ``&NAME ... /`` (also ``$NAME ...
  $END``) with ``KEY=value``, ``KEY(i)=``, ``KEY(i:j)=`` section assignments, repeat
  counts ``r*c`` and ``r*``, null values (``SEMI=,`` or ``RG=,,,`` leave the default,
  returned as ``None``), quoted strings, logicals and E/D exponents.  Names that appear
  without ``=`` (``ZM2,ZMH`` in the reference's INSETUP groups) assign nothing and are
  listed under the group's ``"__bare__"`` key.
* ``read_dat10`` / ``write_dat10``: the NBODY format of KZ(22) = 2, 6, 9, 10 --
  ``READ (10,*) BODY(I), X(1:3,I), XDOT(1:3,I)`` per body (data.F:38-41), with Fortran
  list-directed rules: each READ starts on a new line, takes values from following
  lines while it needs more, and skips what is left of its last line; ``/`` ends the
  READ leaving the remaining items unchanged (zero here).
* ``read_tree``: the TREE format of KZ(22) = 3, 7 (data.F:51-66).
* ``run_parameters``: the values of a parsed file that the block-step driver uses.
"""
from __future__ import annotations

import re
from pathlib import Path

import numpy as np

_NAME = re.compile(r"([A-Za-z][A-Za-z0-9_]*)\s*(\(\s*(\d+)\s*(?::\s*(\d+)\s*)?\))?\s*=")
_INT = re.compile(r"^[+-]?\d+$")
_REAL = re.compile(r"^[+-]?(\d+\.?\d*|\.\d+)([EeDd][+-]?\d+)?$")
_LOGIC = {".TRUE.": True, ".FALSE.": False, "T": True, "F": False, ".T.": True, ".F.": False}


def _strip_comments(text):
    out = []
    for line in text.splitlines():
        q = None
        cut = len(line)
        for k, ch in enumerate(line):
            if q:
                if ch == q:
                    q = None
            elif ch in "'\"":
                q = ch
            elif ch == "!":
                cut = k
                break
        out.append(line[:cut])
    return "\n".join(out)


def _convert(tok):
    t = tok.strip()
    if len(t) >= 2 and t[0] in "'\"" and t[-1] == t[0]:
        return t[1:-1].replace(t[0] * 2, t[0])
    u = t.upper()
    if u in _LOGIC:
        return _LOGIC[u]
    if _INT.match(t):
        return int(t)
    if _REAL.match(t):
        return float(u.replace("D", "E"))
    raise ValueError(f"not a Fortran constant: {tok!r}")


def _values(seg):
    """Value list of one assignment: separators are commas and blanks, an empty field
    between commas is a null value, r*c repeats c, r* repeats null."""
    toks = []
    k, n = 0, len(seg)
    pending_sep = True          # at the start, or after a comma: a value is expected
    while k < n:
        ch = seg[k]
        if ch in " \t\r\n":
            k += 1
            continue
        if ch == ",":
            if pending_sep:
                toks.append(None)
            pending_sep = True
            k += 1
            continue
        if ch in "'\"":
            j = k + 1
            while j < n:
                if seg[j] == ch:
                    if j + 1 < n and seg[j + 1] == ch:
                        j += 2
                        continue
                    break
                j += 1
            toks.append(seg[k:j + 1])
            k = j + 1
            pending_sep = False
            continue
        j = k
        while j < n and seg[j] not in " \t\r\n,":
            j += 1
        toks.append(seg[k:j])
        k = j
        pending_sep = False
    out = []
    for t in toks:
        if t is None:
            out.append(None)
            continue
        m = re.match(r"^(\d+)\*(.*)$", t) if t[0] not in "'\"" else None
        if m:
            r = int(m.group(1))
            out.extend([None if m.group(2) == "" else _convert(m.group(2))] * r)
        else:
            out.append(t)
    return out


def read_namelist(text: str) -> dict:
    """{GROUP: {KEY: value or {index: value}}}; arrays given by section are dicts from
    1-based index to value, scalars are values, null assignments are None."""
    text = _strip_comments(text)
    groups = {}
    for gm in re.finditer(r"[&$]([A-Za-z][A-Za-z0-9_]*)(.*?)(?:/|[&$]END\b)", text, re.S | re.I):
        name = gm.group(1).upper()
        body = gm.group(2)
        g = {}
        bare = []
        marks = list(_NAME.finditer(body))
        for a, m in enumerate(marks):
            key = m.group(1).upper()
            end = marks[a + 1].start() if a + 1 < len(marks) else len(body)
            vals = []
            for t in _values(body[m.end():end]):
                if t is None or not isinstance(t, str):
                    vals.append(t)
                elif re.match(r"^[A-Za-z][A-Za-z0-9_]*$", t) and t.upper() not in _LOGIC:
                    bare.append(t.upper())         # a name without '=': assigns nothing
                else:
                    vals.append(_convert(t))
            if m.group(2):
                lo = int(m.group(3))
                hi = int(m.group(4)) if m.group(4) else lo + max(len(vals), 1) - 1
                arr = g.get(key) if isinstance(g.get(key), dict) else {}
                for q, v in enumerate(vals[:hi - lo + 1]):
                    if v is not None:
                        arr[lo + q] = v
                g[key] = arr
            elif len(vals) <= 1:
                g[key] = vals[0] if vals else None
            else:
                g[key] = {q + 1: v for q, v in enumerate(vals) if v is not None}
        if bare:
            g["__bare__"] = bare
        groups[name] = g
    return groups


def _fmt(v):
    if isinstance(v, bool):
        return ".TRUE." if v else ".FALSE."
    if isinstance(v, str):
        return "'" + v.replace("'", "''") + "'"
    if isinstance(v, float):
        return repr(v).upper()
    return str(v)


def write_namelist(groups: dict) -> str:
    """Inverse of read_namelist (null values are written as empty assignments)."""
    out = []
    for name, g in groups.items():
        parts = []
        for key, v in g.items():
            if key == "__bare__":
                continue
            if isinstance(v, dict):
                for i in sorted(v):
                    parts.append(f"{key}({i})={_fmt(v[i])}")
            elif v is None:
                parts.append(f"{key}=")
            else:
                parts.append(f"{key}={_fmt(v)}")
        out.append(f"&{name}\n" + ",\n".join(parts) + " /\n")
    return "\n".join(out)


def _list_directed(lines, start, need):
    """One list-directed READ of `need` numbers from lines[start:]; returns (values,
    next line).  Missing values after a '/' are zero."""
    vals = []
    k = start
    while len(vals) < need:
        if k >= len(lines):
            raise EOFError(f"end of file after {len(vals)} of {need} values")
        line = lines[k]
        k += 1
        stop = False
        if "/" in line:
            line = line[:line.index("/")]
            stop = True
        for t in _values(line):
            if len(vals) == need:
                break
            vals.append(0.0 if t is None else float(_convert(t)) if isinstance(t, str) else t)
        if stop:
            vals.extend([0.0] * (need - len(vals)))
    return vals[:need], k


def read_dat10(src, n: int | None = None):
    """(BODY[n], X[n][3], XDOT[n][3]) from an NBODY-format file or text; n = all."""
    text = Path(src).read_text() if isinstance(src, (str, Path)) and Path(src).exists() else str(src)
    lines = [ln for ln in text.splitlines()]
    rows = []
    k = 0
    while (n is None or len(rows) < n) and k < len(lines):
        if n is None and not lines[k].strip():
            k += 1
            continue
        v, k = _list_directed(lines, k, 7)
        rows.append(v)
    a = np.array(rows, float).reshape(-1, 7)
    if n is not None and len(a) < n:
        raise EOFError(f"{len(a)} bodies in the file, {n} requested")
    return a[:, 0].copy(), a[:, 1:4].copy(), a[:, 4:7].copy()


def write_dat10(path, m, x, v):
    """One line per body, mass, x, y, z, vx, vy, vz in shortest round-trip form."""
    with open(path, "w") as f:
        for i in range(len(m)):
            f.write(" ".join(repr(float(t)) for t in (m[i], *x[i], *v[i])) + "\n")


def read_tree(src):
    """TREE format (data.F:51-66): N, two dummy lines, N masses, N positions, N
    velocities, one READ each."""
    text = Path(src).read_text() if isinstance(src, (str, Path)) and Path(src).exists() else str(src)
    lines = text.splitlines()
    (nv,), k = _list_directed(lines, 0, 1)
    n = int(nv)
    _, k = _list_directed(lines, k, 1)
    _, k = _list_directed(lines, k, 1)
    m = np.empty(n)
    x = np.empty((n, 3))
    v = np.empty((n, 3))
    for i in range(n):
        (m[i],), k = _list_directed(lines, k, 1)
    for i in range(n):
        x[i], k = _list_directed(lines, k, 3)
    for i in range(n):
        v[i], k = _list_directed(lines, k, 3)
    return m, x, v


def kz(groups, i, default=0):
    """KZ(i) from the ININPUT group."""
    arr = groups.get("ININPUT", {}).get("KZ", {})
    return arr.get(i, default) if isinstance(arr, dict) else default


def run_parameters(groups):
    """Parameters of input.F's ININPUT group that the block-step driver uses."""
    g = groups["ININPUT"]
    get = lambda k, d=None: g.get(k, d) if g.get(k) is not None else d
    return {"n": int(get("N")), "nnbopt": int(get("NNBOPT", 50)), "etai": float(get("ETAI", 0.02)),
            "etar": float(get("ETAR", 0.02)), "rs0": float(get("RS0", 0.3)),
            "smax": float(get("SMAX", 0.125)), "dtmin": float(get("DTMIN", 1e-5)),
            "tcrit": float(get("TCRIT", 10.0)), "nrand": int(get("NRAND", 0)),
            "kz22": int(kz(groups, 22)), "kz39": int(kz(groups, 39)), "kz38": int(kz(groups, 38))}
