"""Host side of the MI355X-native heads-up DCFR postflop solver.

Python mirror of the solver crate's API (crates/solver/src/lib.rs):
`SolveConfig`, `SolveOptions`, `Solver` (new/estimate/solve/query/
ranges_after), `Backend`, over the C ABI of libpts (include/pts.h), whose
solve runs in hand-written HIP kernels for gfx950.
"""
from .config import (BetSize, BettingState, DcfrParams, PostflopMenu, SolveConfig, StreetSizes, make_config,
                     parse_card, parse_cards, card_str, combo_index, combo_cards, FLOP, TURN, RIVER,
                     BACKEND_CPU, BACKEND_GPU, BACKEND_AUTO, STORAGE_F32, STORAGE_COMPRESSED, STORAGE_AUTO)
from .ranges import parse_range, HandClass

__all__ = ["BetSize", "BettingState", "DcfrParams", "PostflopMenu", "SolveConfig", "StreetSizes", "make_config",
           "parse_card", "parse_cards", "card_str", "combo_index", "combo_cards", "parse_range", "HandClass",
           "FLOP", "TURN", "RIVER", "BACKEND_CPU", "BACKEND_GPU", "BACKEND_AUTO", "STORAGE_F32",
           "STORAGE_COMPRESSED", "STORAGE_AUTO"]
