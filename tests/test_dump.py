"""The reference trainer's spot files (POKER_TRAINER_DUMP, crates/trainer/src/
review.rs:1176-1213): the text format, read back into a solve config, and
the same tree built from it by the product and the oracle."""
import numpy as np
import pytest

import pyoracle as O
from pts import PostflopMenu, parse_cards, combo_index
from pts.dump import dump_file_name, menu_debug, parse_menu_debug, read_dump, write_dump
from pts.solver import Solver

# derive(Debug) of the reference's default menu (menu.rs:121-133)
DEFAULT_MENU = ("PostflopMenu { flop: StreetSizes { bet: [Pot(0.33), Pot(0.75)], raise: [Mult(3.0)] }, "
                "turn: StreetSizes { bet: [Pot(0.5), Pot(1.0)], raise: [Mult(3.0)] }, "
                "river: StreetSizes { bet: [Pot(0.5), Pot(1.0)], raise: [Mult(3.0)] }, max_raises: 3, "
                "add_allin_threshold: 1.5, force_allin_threshold: 0.15, merge_threshold: 0.1 }")

# A turn spot as dump_config writes it: player 1 facing a bet of 600 into 1200
# (player 0 bet first), rake 5% capped at 300, no future menu; combos in
# combo-index order, as dump_config loops over them.
TURN_FACING = (
    "label=turn-facing-bet\nboard=Kh7d2cTs\nstreet=Turn\npot=1800\ncommit=600,0\nstack=8400,9000\nto_act=1\n"
    "current_bet=600\nnum_raises=1\nrake=0.05,300\n"
    "menu=PostflopMenu { flop: StreetSizes { bet: [Pot(0.33), Pot(0.75)], raise: [Mult(3.0)] }, "
    "turn: StreetSizes { bet: [Pot(0.5), AllIn], raise: [Mult(2.5), AllIn] }, "
    "river: StreetSizes { bet: [Pot(0.5), Pot(1.0)], raise: [Mult(3.0)] }, max_raises: 2, "
    "add_allin_threshold: 0.0, force_allin_threshold: 0.0, merge_threshold: 0.0 }\n"
    "future_menu=None\n"
    "oop=9s8s:1,QhJh:0.25,KsKd:1,AdAc:0.5,AhAc:1,AhAd:1\n"
    "ip=6c5c:1,8h7h:0.5,JhJd:1,JsJc:0.75,QsJs:1,AsQs:1\n")


def test_default_menu_debug_text():
    assert menu_debug(PostflopMenu()) == DEFAULT_MENU
    assert parse_menu_debug(DEFAULT_MENU) == PostflopMenu()


def test_read_a_dumped_spot():
    label, cfg = read_dump(TURN_FACING)
    assert label == "turn-facing-bet"
    assert cfg.board == parse_cards("Kh7d2cTs")
    r = cfg.root
    assert (r.street, r.pot, r.commit, r.stack, r.to_act, r.current_bet, r.num_raises) == \
        (2, 1800, [600, 0], [8400, 9000], 1, 600, 1)
    # rebuilt: player 1 must act, player 0 already has; the bet is the raise increment
    assert r.needs == [False, True] and r.last_raise_inc == 600
    assert cfg.rake_rate == 0.05 and cfg.rake_cap == 300.0
    assert cfg.future_menu is None and cfg.menu.max_raises == 2
    assert [b.kind for b in cfg.menu.turn.bet] == [0, 2] and cfg.menu.turn.raise_[0].value == 2.5
    w0 = np.asarray(cfg.ranges[0], np.float32)
    assert w0[combo_index(*parse_cards("AdAc"))] == 0.5 and w0[combo_index(*parse_cards("QhJh"))] == 0.25
    assert np.count_nonzero(w0) == 6 and np.count_nonzero(cfg.ranges[1]) == 6
    # writing it back gives the same text
    assert write_dump(cfg, label) == TURN_FACING


def test_round_trip_of_a_street_start_spot():
    label, cfg = read_dump(TURN_FACING)
    cfg.root.commit, cfg.root.current_bet, cfg.root.num_raises, cfg.root.to_act = [0, 0], 0, 0, 0
    cfg.root.pot = 1200
    cfg.future_menu = PostflopMenu()
    text = write_dump(cfg, "street start")
    label2, cfg2 = read_dump(text)
    assert label2 == "street start" and cfg2.future_menu == PostflopMenu()
    assert cfg2.root.needs == [True, True] and cfg2.root.last_raise_inc == 100
    assert write_dump(cfg2, label2) == text


def test_file_names_and_errors():
    assert dump_file_name(7, "turn solve #2") == "hand0007-turn_solve__2.txt"
    with pytest.raises(ValueError):
        read_dump(TURN_FACING.replace("street=Turn", "street=Preflop"))
    with pytest.raises(ValueError):
        read_dump(TURN_FACING.replace("future_menu=None\n", ""))
    with pytest.raises(ValueError):
        parse_menu_debug(DEFAULT_MENU.replace("Mult(3.0)", "Half(3.0)"))


def test_dumped_spot_builds_the_oracle_tree():
    _, cfg = read_dump(TURN_FACING)
    o = O.OracleSolver(cfg)
    s = Solver(cfg, device=-1)
    a, b = o.stats(), s.stats()
    assert (a.player_nodes, a.chance_nodes, a.terminal_nodes, a.storage_cells) == \
        (b.player_nodes, b.chance_nodes, b.terminal_nodes, b.storage_cells)
    assert a.player_nodes > 10
