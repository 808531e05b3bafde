"""Spot files of the reference trainer's `POKER_TRAINER_DUMP` debugging aid
(crates/trainer/src/review.rs:1176-1213, `dump_config`): every heads-up solve
writes its inputs as `key=value` lines, so a real trainer spot can be replayed
here.

    label=<label>
    board=<cards, e.g. Kh7d2c>
    street=<Flop|Turn|River>
    pot=<chips>
    commit=<p0>,<p1>
    stack=<p0>,<p1>
    to_act=<0|1>
    current_bet=<chips>
    num_raises=<n>
    rake=<rate>,<cap>
    menu=<Rust Debug of PostflopMenu>
    future_menu=<None | Some(<Rust Debug of PostflopMenu>)>
    oop=<combo>:<weight>,...   (combo: higher card first, e.g. AhKd:0.5)
    ip=<combo>:<weight>,...

The file does not record the betting state's `last_raise_inc` and `needs`
(tree.rs:17-34); `read_dump` rebuilds them from the recorded fields: the
player to act still has to act; the other player does too only when the
player to act opens the street (nothing bet, player 0 to act); the minimum
raise increment is the amount the player to act faces (at least the big
blind), which is the last raise increment whenever raises grow.
"""
from __future__ import annotations

import os
import re
from typing import Dict, List, Optional, Tuple

import numpy as np

from .config import (COMBOS, FLOP, RIVER, TURN, BetSize, BettingState, PostflopMenu, SolveConfig, StreetSizes,
                     card_str, combo_index, parse_card, parse_cards)

STREET_NAMES = {FLOP: "Flop", TURN: "Turn", RIVER: "River"}
STREET_OF_NAME = {v: k for k, v in STREET_NAMES.items()}


# ------------------------------------------------------------------ numbers
def _f64_debug(x: float) -> str:
    """Rust's `{:?}` of an f64 in the range menus use (1e-4 to 1e16): the
    shortest round-trip digits, with `.0` on integral values (as Python's
    repr)."""
    return repr(float(x))


def _f64_display(x: float) -> str:
    """Rust's `{}` of an f64: shortest round-trip digits, no `.0`."""
    return np.format_float_positional(float(x), unique=True, trim="-")


def _f32_display(x: float) -> str:
    """Rust's `{}` of an f32: the shortest digits that round-trip in f32."""
    return np.format_float_positional(np.float32(x), unique=True, trim="-")


# ------------------------------------------------------------------ menu (Rust Debug)
def _size_debug(b: BetSize) -> str:  # menu.rs:14-24, derive(Debug)
    if b.kind == 0:
        return f"Pot({_f64_debug(b.value)})"
    if b.kind == 1:
        return f"Mult({_f64_debug(b.value)})"
    return "AllIn"


def _street_debug(s: StreetSizes) -> str:  # menu.rs:84-89
    return (f"StreetSizes {{ bet: [{', '.join(_size_debug(b) for b in s.bet)}], "
            f"raise: [{', '.join(_size_debug(b) for b in s.raise_)}] }}")


def menu_debug(m: PostflopMenu) -> str:
    """`format!("{:?}", menu)` of the reference's PostflopMenu (menu.rs:100-119)."""
    return (f"PostflopMenu {{ flop: {_street_debug(m.flop)}, turn: {_street_debug(m.turn)}, "
            f"river: {_street_debug(m.river)}, max_raises: {m.max_raises}, "
            f"add_allin_threshold: {_f64_debug(m.add_allin_threshold)}, "
            f"force_allin_threshold: {_f64_debug(m.force_allin_threshold)}, "
            f"merge_threshold: {_f64_debug(m.merge_threshold)} }}")


_TOKEN = re.compile(r"\s*([A-Za-z_][A-Za-z0-9_]*|-?[0-9][0-9.eE+-]*|[{}\[\](),:])")


def _tokens(text: str) -> List[str]:
    out, i = [], 0
    text = text.strip()
    while i < len(text):
        m = _TOKEN.match(text, i)
        if not m:
            raise ValueError(f"menu: unexpected text at {text[i:i + 20]!r}")
        out.append(m.group(1))
        i = m.end()
    return out


class _Parser:
    def __init__(self, text: str):
        self.t = _tokens(text)
        self.i = 0

    def peek(self) -> Optional[str]:
        return self.t[self.i] if self.i < len(self.t) else None

    def take(self, want: Optional[str] = None) -> str:
        if self.i >= len(self.t):
            raise ValueError("menu: unexpected end")
        x = self.t[self.i]
        if want is not None and x != want:
            raise ValueError(f"menu: expected {want!r}, got {x!r}")
        self.i += 1
        return x

    def size(self) -> BetSize:
        name = self.take()
        if name == "AllIn":
            return BetSize(2, 0.0)
        if name not in ("Pot", "Mult"):
            raise ValueError(f"menu: unknown bet size {name!r}")
        self.take("(")
        v = float(self.take())
        self.take(")")
        return BetSize(0 if name == "Pot" else 1, v)

    def size_list(self) -> List[BetSize]:
        self.take("[")
        out = []
        while self.peek() != "]":
            out.append(self.size())
            if self.peek() == ",":
                self.take(",")
        self.take("]")
        return out

    def struct(self, name: str) -> Dict[str, object]:
        self.take(name)
        self.take("{")
        fields: Dict[str, object] = {}
        while self.peek() != "}":
            key = self.take()
            self.take(":")
            if key in ("flop", "turn", "river"):
                f = self.struct("StreetSizes")
                fields[key] = StreetSizes(f["bet"], f["raise"])
            elif key in ("bet", "raise"):
                fields[key] = self.size_list()
            else:
                fields[key] = self.take()
            if self.peek() == ",":
                self.take(",")
        self.take("}")
        return fields

    def menu(self) -> PostflopMenu:
        f = self.struct("PostflopMenu")
        return PostflopMenu(f["flop"], f["turn"], f["river"], int(f["max_raises"]), float(f["add_allin_threshold"]),
                            float(f["force_allin_threshold"]), float(f["merge_threshold"]))


def parse_menu_debug(text: str) -> PostflopMenu:
    p = _Parser(text)
    m = p.menu()
    if p.peek() is not None:
        raise ValueError("menu: trailing text")
    return m


def parse_future_menu(text: str) -> Optional[PostflopMenu]:
    t = text.strip()
    if t == "None":
        return None
    if not (t.startswith("Some(") and t.endswith(")")):
        raise ValueError(f"future_menu: expected None or Some(...), got {t[:20]!r}")
    return parse_menu_debug(t[5:-1])


# ------------------------------------------------------------------ ranges
def range_text(w) -> str:
    """review.rs:1180-1190: every combo with weight > 0 as <hi><lo>:<weight>,
    in combo order."""
    out = []
    for c in range(1326):
        if w[c] > 0.0:
            a, b = COMBOS[c]  # a has the higher card index, so the higher (or equal) rank
            out.append(f"{card_str(a)}{card_str(b)}:{_f32_display(w[c])}")
    return ",".join(out)


def parse_range_text(text: str) -> List[float]:
    w = [0.0] * 1326
    t = text.strip()
    if not t:
        return w
    for item in t.split(","):
        combo, _, weight = item.strip().partition(":")
        if len(combo) != 4 or not weight:
            raise ValueError(f"range: bad entry {item!r}")
        c = combo_index(parse_card(combo[:2]), parse_card(combo[2:]))
        w[c] = float(np.float32(float(weight)))
    return w


# ------------------------------------------------------------------ files
def write_dump(cfg: SolveConfig, label: str) -> str:
    """The text `dump_config` writes for `cfg` (review.rs:1191-1206)."""
    r = cfg.root
    future = "None" if cfg.future_menu is None else f"Some({menu_debug(cfg.future_menu)})"
    return (f"label={label}\nboard={''.join(card_str(c) for c in cfg.board)}\nstreet={STREET_NAMES[r.street]}\n"
            f"pot={r.pot}\ncommit={r.commit[0]},{r.commit[1]}\nstack={r.stack[0]},{r.stack[1]}\nto_act={r.to_act}\n"
            f"current_bet={r.current_bet}\nnum_raises={r.num_raises}\n"
            f"rake={_f64_display(cfg.rake_rate)},{_f64_display(cfg.rake_cap)}\nmenu={menu_debug(cfg.menu)}\n"
            f"future_menu={future}\noop={range_text(cfg.ranges[0])}\nip={range_text(cfg.ranges[1])}\n")


def dump_file_name(hand: int, label: str) -> str:
    """review.rs:1207-1211: hand<id:04>-<label>.txt, other characters as '_'."""
    name = f"hand{hand:04d}-{label}"
    return "".join(ch if (ch.isascii() and ch.isalnum()) or ch == "-" else "_" for ch in name) + ".txt"


def _pair(v: str) -> List[int]:
    a, b = v.split(",")
    return [int(a), int(b)]


def read_dump(text: str, big_blind: int = 100, **kw) -> Tuple[str, SolveConfig]:
    """(label, SolveConfig) of a dump file's text; keyword arguments set the
    SolveConfig fields the file does not record (dcfr, storage_mode, ...)."""
    f: Dict[str, str] = {}
    for line in text.splitlines():
        if not line.strip():
            continue
        key, sep, val = line.partition("=")
        if not sep:
            raise ValueError(f"dump: not a key=value line: {line[:40]!r}")
        f[key.strip()] = val
    need = ("label", "board", "street", "pot", "commit", "stack", "to_act", "current_bet", "num_raises", "rake",
            "menu", "future_menu", "oop", "ip")
    missing = [k for k in need if k not in f]
    if missing:
        raise ValueError(f"dump: missing {', '.join(missing)}")
    board = parse_cards(f["board"])
    if f["street"] not in STREET_OF_NAME:
        raise ValueError(f"dump: street {f['street']!r} is not a postflop street")
    street = STREET_OF_NAME[f["street"]]
    commit, stack = _pair(f["commit"]), _pair(f["stack"])
    to_act, current_bet = int(f["to_act"]), int(f["current_bet"])
    if to_act not in (0, 1):
        raise ValueError("dump: to_act must be 0 or 1")
    o = to_act ^ 1
    facing = current_bet - commit[to_act]
    needs = [False, False]
    needs[to_act] = stack[to_act] > 0
    needs[o] = current_bet == 0 and to_act == 0 and stack[o] > 0
    inc = max(big_blind, facing) if current_bet > 0 else big_blind
    root = BettingState(street, int(f["pot"]), commit, stack, current_bet, inc, int(f["num_raises"]), to_act, needs)
    rake = f["rake"].split(",")
    cfg = SolveConfig(board, [parse_range_text(f["oop"]), parse_range_text(f["ip"])], root,
                      parse_menu_debug(f["menu"]), future_menu=parse_future_menu(f["future_menu"]),
                      big_blind=big_blind, rake_rate=float(rake[0]), rake_cap=float(rake[1]), **kw)
    return f["label"], cfg


def load_dir(path: str, **kw) -> List[Tuple[str, str, SolveConfig]]:
    """Every dump file of a POKER_TRAINER_DUMP folder: (file, label, config),
    in file-name order."""
    out = []
    for name in sorted(os.listdir(path)):
        if name.endswith(".txt"):
            with open(os.path.join(path, name)) as fh:
                label, cfg = read_dump(fh.read(), **kw)
            out.append((name, label, cfg))
    return out
