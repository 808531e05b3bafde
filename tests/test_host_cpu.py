"""Host-side tests that need no GPU: the C ABI library loads and exports
every declared symbol, the product's tree arena equals the oracle's (the
storage layout is the boundary format), the plan validator rejects corrupted
trees before anything reaches the device (as crates/solver/src/gpu/plan/
tests.rs:158-248 does for the reference's plan)."""
import ctypes
import os
import re

import numpy as np
import pytest

import pyoracle as O
from pts import native, spots, PostflopMenu, make_config
from pts.config import ConfigC
from pts.native import lib, check, PtsError, TreeDescC
from pts.solver import TreeArrays

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    L = lib()
    text = open(os.path.join(ROOT, "include", "pts.h")).read()
    names = set(re.findall(r"\b(pts_[a-z_0-9]+)\s*\(", text))
    assert len(names) >= 30
    for n in sorted(names):
        assert hasattr(L, n), n
    assert set(native.EXPORTS) <= names


def test_struct_layouts_match_the_header():
    # The one ctypes layout serves pts_config_c and the oracle's orc_config_c.
    from pts.native import ProgressC, StepC, StatsC
    L = lib()
    for what, ct in enumerate((ConfigC, TreeDescC, ProgressC, StepC, StatsC)):
        assert L.pts_abi_sizeof(what) == ctypes.sizeof(ct), ct
    O.lib().orc_config_size.restype = ctypes.c_size_t
    assert O.lib().orc_config_size() == ctypes.sizeof(ConfigC)


def host_solver(cfg):
    c = cfg.to_c()
    h = ctypes.c_void_p()
    check(lib().pts_solver_new(ctypes.byref(c), -1, 1, 0, ctypes.byref(h)))
    return h


def product_nodes(h):
    m = lib().pts_solver_num_nodes(h)
    a = dict(kind=np.zeros(m, np.uint8), street=np.zeros(m, np.uint8), num_children=np.zeros(m, np.uint16),
             first_child=np.zeros(m, np.uint32), invest=np.zeros((m, 2), np.int32), offset=np.zeros(m, np.uint64),
             rake=np.zeros(m, np.float32), chance_factor=np.zeros(m, np.float32), card=np.zeros(m, np.int8),
             edge_act=np.zeros(m, np.int32), edge_to=np.zeros(m, np.int64))
    lib().pts_solver_nodes(h, *[x.ctypes.data for x in a.values()])
    return a


def iso_turn():
    # diamonds and spades are absent from the board: every river card of one
    # of them is merged onto the other's
    return make_config("2c7c9hTh", "AA, KK, QQ, AKs, AQs, KQs, JTs, 98s, 76s", "TT, 99, 88, AJs, KJs, QJs, T9s, 87s, 65s",
                       1000, 5000, PostflopMenu.one(["50%"], [], max_raises=1), isomorphism=True)


def iso_flop():
    # one card of three suits: no merge at the turn; after a turn card that
    # repeats a rank in spades, that pair of suits merges river cards
    return make_config("2c7d9h", "AA, KK, QQ, AKs, JTs, 98s", "TT, 99, AJs, KQs, 87s, 65s", 600, 1200,
                       PostflopMenu.one(["50%"], [], max_raises=1), isomorphism=True)


CONFIGS = [
    ("river", lambda: spots.river_spot(0)),
    ("river7", lambda: spots.river_spot(7)),
    ("turn", lambda: spots.turn_spot(0)),
    ("toy", lambda: make_config("Kc8d5h3s2c", "AA, JTs, J9s:0.5", "KQ", 1000, 5000, PostflopMenu.one(["100%"], []))),
    ("allin", lambda: make_config("2c7d9hTs", "AhAd", "KhKd", 1000, 1000, PostflopMenu.one(["allin"], []))),
    ("default-menu", lambda: make_config("Qh9d4c2s", "AQ, KQ, QJs, JTs, 99, 44, T9s, 87s, A5s, KJs",
                                         "AA, KK, AQ, QT, 99, JJ, TT, 98s, KJs, 65s, AJs", 600, 4000, PostflopMenu())),
    # suit isomorphism (iso.rs): class ranges are suit-symmetric, and the
    # board leaves suits with equal rank sets, so chance nodes merge cards
    ("iso-turn", lambda: iso_turn()),
    ("iso-flop", lambda: iso_flop()),
]


@pytest.mark.parametrize("name,make", CONFIGS, ids=[c[0] for c in CONFIGS])
def test_tree_arena_equals_the_oracle(name, make):
    cfg = make()
    h = host_solver(cfg)
    try:
        a = product_nodes(h)
        o = O.OracleSolver(cfg)
        b = o.nodes()
        assert len(a["kind"]) == len(b["kind"])
        for k in ("kind", "street", "num_children", "first_child", "invest", "rake", "chance_factor"):
            assert np.array_equal(a[k], b[k]), k
        player = a["kind"] <= 1
        assert np.array_equal(a["offset"][player], b["offset"][player])
        sd = a["kind"] == 5
        # showdown tables: same board per table index
        ob = o.boards()
        assert np.array_equal(a["offset"][sd], b["offset"][sd])
        assert len(ob) == (int(a["offset"][sd].max()) + 1 if sd.any() else 0)
        ch = b["edge_card"] >= 0
        assert np.array_equal(a["card"][ch].astype(np.int32), b["edge_card"][ch])
        act = b["edge_act"] >= 0
        assert np.array_equal(a["edge_act"][act], b["edge_act"][act])
        assert np.array_equal(a["edge_to"][act], b["edge_to"][act])
        st = O.OracleSolver.estimate(cfg)
        assert lib().pts_solver_storage_len(h) == o.storage_len == st.storage_cells
    finally:
        lib().pts_solver_free(h)


def test_estimate_matches_the_oracle():
    for make in (lambda: spots.turn_spot(1), lambda: spots.river_spot(2), iso_turn, iso_flop):
        cfg = make()
        from pts.solver import Solver
        a = Solver.estimate(cfg)
        b = O.OracleSolver.estimate(cfg)
        assert (a.player_nodes, a.chance_nodes, a.terminal_nodes, a.storage_cells, a.memory_bytes) == \
            (b.player_nodes, b.chance_nodes, b.terminal_nodes, b.storage_cells, b.memory_bytes)


def test_solve_config_new_takes_the_reference_defaults():
    """lib.rs:107-134: automatic storage, isomorphism on, 8 GB, no rake.
    Automatic storage keeps f32 (8 bytes a cell) under 1 GB of f32 tree and
    takes 16-bit storage (4 bytes a cell) above it (lib.rs:346-352,
    tree.rs:264-272); every node costs 56 bytes either way."""
    from pts import SolveConfig, STORAGE_AUTO
    from pts.solver import Solver
    for make, cell_bytes in ((lambda: spots.river_spot(2), 8), (lambda: spots.flop_spot(0), 4)):
        c = make()
        n = SolveConfig.new(c.board, c.ranges, c.root, c.menu)
        assert n.storage_mode == STORAGE_AUTO and n.isomorphism
        assert (n.max_memory_bytes, n.big_blind, n.prune, n.rake_rate, n.rake_cap) == (8 << 30, 100, 0.0, 0.0, 0.0)
        assert n.future_menu is None and n.sampled is None
        d = n.dcfr  # cfr.rs:65-81
        assert (d.alpha, d.beta, d.gamma, d.restart_average, d.unweighted_average) == (1.5, 0.0, 3.0, True, True)
        st = Solver.estimate(n)
        nodes = st.player_nodes + st.chance_nodes + st.terminal_nodes
        assert st.memory_bytes == st.storage_cells * cell_bytes + nodes * 56
        assert (st.storage_cells * 8 + nodes * 56 > (1 << 30)) == (cell_bytes == 4)


def test_rejections():
    cfg = spots.river_spot(0)
    cfg.max_memory_bytes = 1 << 10
    with pytest.raises(PtsError, match="GB but the limit"):
        host_solver(cfg)
    cfg = spots.river_spot(0)
    cfg.board = cfg.board[:4]
    with pytest.raises(PtsError, match="board"):
        host_solver(cfg)
    cfg = spots.river_spot(0)
    cfg.ranges[0] = [0.0] * 1326
    with pytest.raises(PtsError, match="empty range"):
        host_solver(cfg)
    # class-based (suit-symmetric) ranges with isomorphism requested
    cfg = spots.flop_spot(0)
    cfg.isomorphism = True
    cfg.max_memory_bytes = 1 << 20
    with pytest.raises(PtsError, match="isomorphism|limit"):
        host_solver(cfg)


def oracle_tree(cfg):
    o = O.OracleSolver(cfg)
    b = o.nodes()
    hands = [o.hands(p) for p in range(2)]
    from pts.config import COMBOS
    cards = [np.array([COMBOS[int(c)] for c in hands[p][0]], np.uint8) for p in range(2)]
    lookup = [{int(c): i for i, c in enumerate(hands[p][0])} for p in range(2)]
    same = [np.array([lookup[p ^ 1].get(int(c), 0xFFFFFFFF) for c in hands[p][0]], np.uint32) for p in range(2)]
    card = np.where(b["edge_card"] >= 0, b["edge_card"], -1).astype(np.int8)
    boards = o.boards()
    t = TreeArrays(b["kind"], b["street"], b["num_children"], b["first_child"], b["invest"], b["offset"], b["rake"],
                   b["chance_factor"], card, boards, cards, [hands[p][1] for p in range(2)], same,
                   cfg.root.pot / 2.0, o.num_combinations, o.storage_len)
    return o, t


def create_status(t, n):
    h = ctypes.c_void_p()
    z = np.zeros(n, np.float32)
    st = lib().pts_engine_create(0, ctypes.byref(t.desc), z.ctypes.data, z.ctypes.data, 1, 0, ctypes.byref(h))
    if st == 0:
        lib().pts_engine_destroy(h)
    return st


MUTATIONS = [
    ("kind", lambda a: a.__setitem__(3, 9)),
    ("first_child", lambda a: a.__setitem__(0, 0)),
    ("first_child_far", lambda a: a.__setitem__(0, 10 ** 8)),
    ("num_children", lambda a: a.__setitem__(0, 0)),
    ("offset", lambda a: a.__setitem__(0, 10 ** 9)),
    ("chance_factor", None),
    ("card", None),
]


@pytest.mark.parametrize("field", ["kind", "first_child", "first_child_far", "num_children", "offset", "card",
                                   "chance_factor", "hand_cards", "hand_same", "boards", "storage_len"])
def test_plan_rejects_corrupted_trees(field):
    _, t = oracle_tree(spots.turn_spot(0) if field in ("card", "chance_factor") else spots.river_spot(0))
    a = t.a
    if field == "kind":
        a["kind"][5] = 9
    elif field == "first_child":
        a["first_child"][0] = 0
    elif field == "first_child_far":
        a["first_child"][0] = 10 ** 8
    elif field == "num_children":
        a["num_children"][0] = 0
    elif field == "offset":
        a["offset"][0] = 10 ** 9
    elif field == "card":
        ch = np.nonzero(a["card"] >= 0)[0]
        a["card"][ch[0]] = -1
    elif field == "chance_factor":
        c = np.nonzero(a["kind"] == 2)[0]
        a["chance_factor"][c[0]] = 0.0
    elif field == "hand_cards":
        t.hc[0][0] = 60
    elif field == "hand_same":
        t.hs[0][0] = 5000
    elif field == "boards":
        a["boards"][1] = a["boards"][0]
    elif field == "storage_len":
        t.desc.storage_len = 10
    st = create_status(t, max(int(t.desc.storage_len), 1))
    assert st == 5, native.last_error()  # PTS_INVALID_INPUT, before any device work
    assert "invalid" in native.last_error()


def test_valid_tree_passes_validation_without_a_gpu():
    if native.device_count() > 0:
        pytest.skip("a GPU is present: creation would succeed")
    _, t = oracle_tree(spots.river_spot(0))
    st = create_status(t, int(t.desc.storage_len))
    assert st == 3  # PTS_NO_DEVICE: validation passed, the device is missing


def test_error_details():
    # gpu/mod.rs:86-136: the numbers behind Driver and OutOfMemory errors;
    # a validation error carries none
    _, t = oracle_tree(spots.river_spot(0))
    t.desc.storage_len = 10
    assert create_status(t, 10) == 5
    assert native.last_error_detail() == (0, 0, 0)
    if native.device_count() == 0:
        with pytest.raises(native.PtsError) as e:
            native.check(lib().pts_device_count(ctypes.byref(ctypes.c_int())))
        assert e.value.status == 3 and e.value.hip_error != 0  # the HIP error of the failed device query


def test_backend_switch():
    # lib.rs:497-503, with Auto set from the MI355X measurement (bench.py
    # --workload river-auto): every street on the GPU
    L = lib()
    assert L.pts_backend_wants_gpu(0, 2) == 0 and L.pts_backend_wants_gpu(0, 3) == 0
    assert L.pts_backend_wants_gpu(1, 3) == 1
    assert all(L.pts_backend_wants_gpu(2, s) == 1 for s in (1, 2, 3))
    assert L.pts_backend_wants_gpu(2, 0) == 0 and L.pts_backend_wants_gpu(2, 4) == 0


def test_spot_generators_are_deterministic():
    a, b = spots.turn_spot(5), spots.turn_spot(5)
    assert a.board == b.board and a.ranges == b.ranges
    assert spots.turn_spot(5).board != spots.turn_spot(6).board
    r = spots.river_spot(0)
    nz = sum(1 for x in r.ranges[0] if x > 0)
    assert 500 < nz < 800  # 1081 unblocked combos, 40% zeroed


# ---------------------------------------------------------------- chunk groups
def plan_groups(h, budget, f64):
    n, b = ctypes.c_uint32(), ctypes.c_uint64()
    check(lib().pts_solver_plan_groups(h, budget, f64, ctypes.byref(n), ctypes.byref(b)))
    return n.value, b.value


@pytest.mark.parametrize("name,make", [("river", lambda: spots.river_spot(0)), ("turn", lambda: spots.turn_spot(0)),
                                       ("flop", lambda: spots.small_flop_spot(0))])
def test_chunk_groups_map_every_row_into_the_scratch(name, make):
    # Every reach slot and value row a grouped traversal touches maps inside
    # the scratch the groups were sized for, at every budget (the check runs
    # on the host before an engine would launch anything).
    h = host_solver(make())
    try:
        for f64 in (0, 1):
            one, full = plan_groups(h, 1 << 62, f64)
            assert one == 1
            most, least = plan_groups(h, 0, f64)
            assert least <= full
            if name == "river":
                assert most == 1 and least == full
                continue
            assert most > 1
            for budget in (full // 7, full // 3, full // 2):
                n, b = plan_groups(h, budget, f64)
                assert 1 <= n <= most
                assert b <= max(budget, least)
    finally:
        lib().pts_solver_free(h)


def test_flop_plan_has_one_chunk_per_turn_subtree():
    cfg = spots.small_flop_spot(0)
    h = host_solver(cfg)
    try:
        most, _ = plan_groups(h, 0, 0)
        nodes = product_nodes(h)
        # chunks are the chance children of the flop's chance nodes
        kind, fc, nc = nodes["kind"], nodes["first_child"], nodes["num_children"]
        street = nodes["street"]
        flop_chance = [i for i in range(len(kind)) if kind[i] == 2 and street[i] == 1]  # street 1: flop
        assert most == sum(int(nc[i]) for i in flop_chance)
    finally:
        lib().pts_solver_free(h)
