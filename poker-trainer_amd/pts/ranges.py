"""Range notation -> 1326 weights (crates/holdem/src/range.rs:201-264) and
the 169 hand classes (crates/holdem/src/combo.rs:56-207)."""
from __future__ import annotations

from typing import List

from .config import RANKS, combo_index, parse_cards


class HandClass:
    def __init__(self, r1: int, r2: int, suited: bool):
        self.hi, self.lo = (r1, r2) if r1 >= r2 else (r2, r1)
        self.suited = suited and self.hi != self.lo

    def grid_index(self) -> int:  # combo.rs:90-97
        rh, rl = 12 - self.hi, 12 - self.lo
        return rh * 13 + rl if self.suited else rl * 13 + rh

    @staticmethod
    def from_grid_index(g: int) -> "HandClass":  # combo.rs:99-109
        row, col = divmod(g, 13)
        r1, r2 = 12 - row, 12 - col
        return HandClass(r1, r2, row < col)

    def combos(self) -> List[int]:  # combo.rs:127-149
        out = []
        for s1 in range(4):
            for s2 in range(4):
                a, b = self.hi * 4 + s1, self.lo * 4 + s2
                if a == b:
                    continue
                if self.hi == self.lo:
                    ok = s1 < s2
                elif self.suited:
                    ok = s1 == s2
                else:
                    ok = s1 != s2
                if ok:
                    out.append(combo_index(a, b))
        return out


def _pattern(s: str):
    if not 2 <= len(s) <= 3:
        raise ValueError(s)
    r1, r2 = RANKS.find(s[0].upper()), RANKS.find(s[1].upper())
    if r1 < 0 or r2 < 0:
        raise ValueError(s)
    hi, lo = max(r1, r2), min(r1, r2)
    if len(s) == 2:
        return hi, lo, None
    c = s[2].lower()
    if c == "s" and hi != lo:
        return hi, lo, True
    if c == "o" and hi != lo:
        return hi, lo, False
    raise ValueError(s)


def _classes(p):
    hi, lo, suited = p
    if hi == lo:
        return [HandClass(hi, lo, False)]
    if suited is not None:
        return [HandClass(hi, lo, suited)]
    return [HandClass(hi, lo, True), HandClass(hi, lo, False)]


def _expand(tok: str) -> List[int]:
    low = tok.lower()
    if low in ("any", "random", "all"):
        return list(range(1326))
    if len(tok) == 4:
        try:
            cs = parse_cards(tok)
        except ValueError:
            cs = None
        if cs:
            if cs[0] == cs[1]:
                raise ValueError(tok)
            return [combo_index(cs[0], cs[1])]
    pats = []
    if "-" in tok:
        a, b = tok.split("-", 1)
        pa, pb = _pattern(a.strip()), _pattern(b.strip())
        if pa[0] == pa[1] and pb[0] == pb[1]:
            pats = [(r, r, None) for r in range(min(pa[0], pb[0]), max(pa[0], pb[0]) + 1)]
        elif pa[0] == pb[0] and pa[2] == pb[2] and pa[0] != pa[1] and pb[0] != pb[1]:
            pats = [(pa[0], l, pa[2]) for l in range(min(pa[1], pb[1]), max(pa[1], pb[1]) + 1)]
        else:
            raise ValueError(f"invalid span: {tok!r}")
    elif tok.endswith("+"):
        p = _pattern(tok[:-1].strip())
        if p[0] == p[1]:
            pats = [(r, r, None) for r in range(p[0], 13)]
        else:
            pats = [(p[0], l, p[2]) for l in range(p[1], p[0])]
    else:
        pats = [_pattern(tok)]
    out = []
    for p in pats:
        for c in _classes(p):
            out.extend(c.combos())
    return out


def parse_range(s: str) -> List[float]:
    w = [0.0] * 1326
    for raw in s.split(","):
        raw = raw.strip()
        if not raw:
            continue
        if ":" in raw:
            t, x = raw.split(":", 1)
            tok, weight = t.strip(), float(x.strip())
        else:
            tok, weight = raw, 1.0
        for c in _expand(tok):
            w[c] = weight
    return w
