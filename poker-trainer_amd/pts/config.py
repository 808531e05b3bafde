"""Solve configuration: the Python mirror of the solver crate's `SolveConfig`,
`PostflopMenu`, `BettingState` and `DcfrParams`.

Reference: crates/solver/src/lib.rs:73-135 (SolveConfig), crates/holdem/src/
menu.rs:14-133 (BetSize, StreetSizes, PostflopMenu), crates/solver/src/
tree.rs:17-61 (BettingState), crates/solver/src/cfr.rs:47-81 (DcfrParams).

`to_c()` packs a config into `pts_config_c` (include/pts.h), the flat struct
the C ABI takes.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field, replace
from typing import List, Optional

RANKS = "23456789TJQKA"
SUITS = "cdhs"

FLOP, TURN, RIVER = 1, 2, 3
STREET_OF_BOARD = {3: FLOP, 4: TURN, 5: RIVER}


def parse_card(s: str) -> int:
    """crates/holdem/src/card.rs:113-126: rank*4 + suit."""
    s = s.strip()
    if len(s) != 2:
        raise ValueError(f"invalid card: {s!r}")
    r = RANKS.find(s[0].upper())
    u = SUITS.find(s[1].lower())
    if r < 0 or u < 0:
        raise ValueError(f"invalid card: {s!r}")
    return r * 4 + u


def parse_cards(s: str) -> List[int]:
    """card.rs:129-136: "AhKd7c" or "Ah Kd 7c"."""
    c = "".join(ch for ch in s if not ch.isspace() and ch != ",")
    if len(c) % 2:
        raise ValueError(f"invalid card list: {s!r}")
    return [parse_card(c[i:i + 2]) for i in range(0, len(c), 2)]


def card_str(c: int) -> str:
    return RANKS[c >> 2] + SUITS[c & 3]


def combo_index(a: int, b: int) -> int:
    """crates/holdem/src/combo.rs:12-17."""
    lo, hi = (a, b) if a < b else (b, a)
    return hi * (hi - 1) // 2 + lo


def combo_cards(i: int):
    """combo.rs:19-41: (higher-index card, lower-index card)."""
    hi = 1
    while (hi + 1) * hi // 2 <= i:
        hi += 1
    return hi, i - hi * (hi - 1) // 2


COMBOS = [combo_cards(i) for i in range(1326)]


@dataclass(frozen=True)
class BetSize:
    """menu.rs:14-51. kind 0 = pot fraction, 1 = multiple of the bet, 2 = all-in."""
    kind: int
    value: float = 0.0

    @staticmethod
    def parse(s: str) -> "BetSize":
        t = s.strip().lower()
        if t in ("a", "allin", "all-in", "all in", "shove"):
            return BetSize(2, 0.0)
        if t.endswith("%"):
            v = float(t[:-1])
            if v <= 0:
                raise ValueError(s)
            return BetSize(0, v / 100.0)
        if t.endswith("x"):
            v = float(t[:-1])
            if v <= 1.0:
                raise ValueError(s)
            return BetSize(1, v)
        raise ValueError(f"invalid bet size {s!r}")


def sizes(*xs: str) -> List[BetSize]:
    return [BetSize.parse(x) for x in xs]


@dataclass
class StreetSizes:
    bet: List[BetSize]
    raise_: List[BetSize]


@dataclass
class PostflopMenu:
    """menu.rs:100-133, defaults as in the reference."""
    flop: StreetSizes = field(default_factory=lambda: StreetSizes(sizes("33%", "75%"), sizes("3x")))
    turn: StreetSizes = field(default_factory=lambda: StreetSizes(sizes("50%", "100%"), sizes("3x")))
    river: StreetSizes = field(default_factory=lambda: StreetSizes(sizes("50%", "100%"), sizes("3x")))
    max_raises: int = 3
    add_allin_threshold: float = 1.5
    force_allin_threshold: float = 0.15
    merge_threshold: float = 0.1

    @staticmethod
    def one(bet, raise_, max_raises: int = 3) -> "PostflopMenu":
        """The explicit menu of the reference's tests (tests.rs:6-17): the
        same sizes on every street and no all-in add/force/merge rules."""
        s = StreetSizes(sizes(*bet), sizes(*raise_))
        return PostflopMenu(s, replace(s), replace(s), max_raises, 0.0, 0.0, 0.0)


@dataclass
class BettingState:
    """tree.rs:17-61. Player 0 is out of position."""
    street: int
    pot: int
    commit: List[int]
    stack: List[int]
    current_bet: int
    last_raise_inc: int
    num_raises: int
    to_act: int
    needs: List[bool]

    @staticmethod
    def street_start(street: int, pot: int, stack0: int, stack1: int, big_blind: int = 100) -> "BettingState":
        return BettingState(street, pot, [0, 0], [stack0, stack1], 0, big_blind, 0, 0, [stack0 > 0, stack1 > 0])


@dataclass
class DcfrParams:
    """cfr.rs:47-81. The reference default is gamma 3 with average restarts
    and unweighted sums; the workloads here use textbook DCFR
    (alpha 1.5, beta 0, gamma 2, reach-weighted sums, no restarts)."""
    alpha: float = 1.5
    beta: Optional[float] = 0.0
    gamma: float = 2.0
    restart_average: bool = False
    unweighted_average: bool = False

    @staticmethod
    def reference_default() -> "DcfrParams":
        return DcfrParams(1.5, 0.0, 3.0, True, True)


BACKEND_CPU, BACKEND_GPU, BACKEND_AUTO = 0, 1, 2

# storage.rs:80-85: f32, 16-bit ("compressed"), automatic (16-bit when the f32
# tree would exceed 1 GB).
STORAGE_F32, STORAGE_COMPRESSED, STORAGE_AUTO = 0, 1, 2


@dataclass
class SolveConfig:
    """lib.rs:73-135.

    The field defaults ask for f32 storage and no isomorphism, the plainest
    tree. `SolveConfig.new` gives the defaults of the reference's
    `SolveConfig::new` (lib.rs:107-134): automatic storage, isomorphism on and
    the reference's default DCFR parameters.
    """
    board: List[int]
    ranges: List[List[float]]
    root: BettingState
    menu: PostflopMenu
    future_menu: Optional[PostflopMenu] = None
    big_blind: int = 100
    prune: float = 0.0
    max_memory_bytes: int = 8 << 30
    rake_rate: float = 0.0
    rake_cap: float = 0.0
    storage_mode: int = 0
    isomorphism: bool = False
    dcfr: DcfrParams = field(default_factory=DcfrParams)
    backend: int = BACKEND_GPU
    # tree.rs:128-136: None (Exact) or (turn cards, river cards, seed) sampled per branch
    sampled: Optional[tuple] = None

    @classmethod
    def new(cls, board: List[int], ranges: List[List[float]], root: BettingState,
            menu: PostflopMenu) -> "SolveConfig":
        """lib.rs:107-134: exact runouts, no rake, automatic storage,
        isomorphism on, an 8 GB limit, the reference's default DCFR
        parameters (cfr.rs:65-81: gamma 3, unweighted sums, restarts); the
        solve runs on the GPU."""
        return cls(board=list(board), ranges=[list(r) for r in ranges], root=root, menu=menu,
                   storage_mode=STORAGE_AUTO, isomorphism=True, dcfr=DcfrParams.reference_default())

    def to_c(self) -> "ConfigC":
        c = ConfigC()
        for i, x in enumerate(self.board):
            c.board[i] = x
        c.board_len = len(self.board)
        for p in range(2):
            arr = (ctypes.c_float * 1326)(*[float(x) for x in self.ranges[p]])
            ctypes.memmove(ctypes.addressof(c.ranges) + p * 1326 * 4, arr, 1326 * 4)
        r = self.root
        c.street, c.to_act, c.pot = r.street, r.to_act, r.pot
        c.commit[0], c.commit[1] = r.commit
        c.stack[0], c.stack[1] = r.stack
        c.current_bet, c.last_raise_inc, c.num_raises = r.current_bet, r.last_raise_inc, r.num_raises
        c.needs[0], c.needs[1] = int(r.needs[0]), int(r.needs[1])
        _menu_to_c(self.menu, c.menu)
        c.has_future_menu = int(self.future_menu is not None)
        if self.future_menu is not None:
            _menu_to_c(self.future_menu, c.future_menu)
        c.big_blind = self.big_blind
        c.rake_rate, c.rake_cap = self.rake_rate, self.rake_cap
        c.max_memory_bytes = self.max_memory_bytes
        c.prune = self.prune
        c.storage_mode = self.storage_mode
        c.isomorphism = int(self.isomorphism)
        d = self.dcfr
        c.beta_some = int(d.beta is not None)
        c.alpha, c.beta, c.gamma = d.alpha, (d.beta or 0.0), d.gamma
        c.restart_average, c.unweighted_average = int(d.restart_average), int(d.unweighted_average)
        c.backend = self.backend
        if self.sampled is not None:
            c.chance_mode = 1
            c.sample_turn, c.sample_river, c.sample_seed = int(self.sampled[0]), int(self.sampled[1]), int(self.sampled[2])
        return c


class SizesC(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32 * 8), ("value", ctypes.c_double * 8), ("count", ctypes.c_int32),
                ("_pad", ctypes.c_int32)]


class MenuC(ctypes.Structure):
    _fields_ = [("bet", SizesC * 3), ("raise_", SizesC * 3), ("max_raises", ctypes.c_uint32), ("_pad", ctypes.c_uint32),
                ("add_allin_threshold", ctypes.c_double), ("force_allin_threshold", ctypes.c_double),
                ("merge_threshold", ctypes.c_double)]


class ConfigC(ctypes.Structure):
    """pts_config_c (include/pts.h) == orc_config_c (the oracle's copy)."""
    _fields_ = [
        ("board", ctypes.c_int32 * 5), ("board_len", ctypes.c_int32),
        ("ranges", (ctypes.c_float * 1326) * 2),
        ("street", ctypes.c_int32), ("to_act", ctypes.c_int32),
        ("pot", ctypes.c_int64), ("commit", ctypes.c_int64 * 2), ("stack", ctypes.c_int64 * 2),
        ("current_bet", ctypes.c_int64), ("last_raise_inc", ctypes.c_int64),
        ("num_raises", ctypes.c_uint32), ("needs", ctypes.c_int32 * 2), ("has_future_menu", ctypes.c_int32),
        ("menu", MenuC), ("future_menu", MenuC),
        ("big_blind", ctypes.c_int64), ("rake_rate", ctypes.c_double), ("rake_cap", ctypes.c_double),
        ("max_memory_bytes", ctypes.c_uint64), ("prune", ctypes.c_float), ("storage_mode", ctypes.c_int32),
        ("isomorphism", ctypes.c_int32), ("beta_some", ctypes.c_int32),
        ("alpha", ctypes.c_double), ("beta", ctypes.c_double), ("gamma", ctypes.c_double),
        ("restart_average", ctypes.c_int32), ("unweighted_average", ctypes.c_int32),
        ("backend", ctypes.c_int32), ("chance_mode", ctypes.c_int32),
        ("sample_turn", ctypes.c_uint32), ("sample_river", ctypes.c_uint32), ("sample_seed", ctypes.c_uint64),
    ]


def _sizes_to_c(lst: List[BetSize], out: SizesC):
    if len(lst) > 8:
        raise ValueError("at most 8 sizes per list")
    for i, b in enumerate(lst):
        out.kind[i] = b.kind
        out.value[i] = b.value
    out.count = len(lst)


def _menu_to_c(m: PostflopMenu, out: MenuC):
    for i, st in enumerate((m.flop, m.turn, m.river)):
        _sizes_to_c(st.bet, out.bet[i])
        _sizes_to_c(st.raise_, out.raise_[i])
    out.max_raises = m.max_raises
    out.add_allin_threshold = m.add_allin_threshold
    out.force_allin_threshold = m.force_allin_threshold
    out.merge_threshold = m.merge_threshold


def make_config(board: str, oop, ip, pot: int, stack: int, menu: PostflopMenu, **kw) -> SolveConfig:
    """The reference tests' `config` helper (crates/solver/src/tests.rs:19-28):
    ranges are notation strings or 1326-weight lists. It keeps the field
    defaults (f32 storage, isomorphism off) unless `kw` asks otherwise; the
    reference helper builds through `SolveConfig::new`, whose isomorphism is
    on, so a test that follows one of its cases passes `isomorphism=True`."""
    from .ranges import parse_range
    cards = parse_cards(board)
    street = STREET_OF_BOARD[len(cards)]
    r = [parse_range(x) if isinstance(x, str) else list(x) for x in (oop, ip)]
    return SolveConfig(cards, r, BettingState.street_start(street, pot, stack, stack), menu, **kw)
