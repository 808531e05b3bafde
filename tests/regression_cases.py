"""The reference's proptest regression inputs (tests/golden/regressions.json,
written by tests/golden/make_regressions.py) as SolveConfigs."""
import json
import os

from pts import BettingState, DcfrParams, PostflopMenu
from pts.config import BetSize, SolveConfig, StreetSizes

HERE = os.path.dirname(os.path.abspath(__file__))


def load():
    return json.load(open(os.path.join(HERE, "golden", "regressions.json")))


def _sizes(xs):
    return [BetSize(int(k), float(v)) for k, v in xs]


def engine_config(c) -> SolveConfig:
    """A shrunk SolveConfig of crates/solver/src/gpu/engine/tests.rs, as stored."""
    m = c["menu"]
    menu = PostflopMenu(*[StreetSizes(_sizes(m[s]["bet"]), _sizes(m[s]["raise"])) for s in ("flop", "turn", "river")],
                        m["max_raises"], m["add_allin_threshold"], m["force_allin_threshold"], m["merge_threshold"])
    r = c["root"]
    root = BettingState(r["street"], r["pot"], r["commit"], r["stack"], r["current_bet"], r["last_raise_inc"],
                        r["num_raises"], r["to_act"], r["needs"])
    d = c["dcfr"]
    return SolveConfig(c["board"], c["ranges"], root, menu, rake_rate=c["rake_rate"], rake_cap=c["rake_cap"],
                       isomorphism=c["isomorphism"],
                       dcfr=DcfrParams(d["alpha"], d["beta"], d["gamma"], d["restart_average"], d["unweighted_average"]))


def river_config(c) -> SolveConfig:
    """crates/solver/src/proptests.rs:186-201 (`river_config`, two sizes and a
    raise) on a stored river case; a case stored without pot / stack (a
    showdown property's inputs) uses pot 1000 and stacks 2000."""
    pot, stack = c["pot"] or 1000, c["stack"] or 2000
    s = StreetSizes([BetSize(0, 0.5), BetSize(0, 1.0)], [BetSize(1, 3.0)])
    menu = PostflopMenu(s, StreetSizes(list(s.bet), list(s.raise_)), StreetSizes(list(s.bet), list(s.raise_)))
    return SolveConfig(c["board"], c["ranges"], BettingState.street_start(3, pot, stack, stack), menu,
                       isomorphism=True, dcfr=DcfrParams.reference_default())


def all_configs():
    d = load()
    out = [("engine-" + c["id"], engine_config(c)) for c in d["engine"]]
    out += [("river-" + c["id"], river_config(c)) for c in d["river"]]
    return out
