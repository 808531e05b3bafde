"""Writes tests/golden/regressions.json from the reference's proptest
regression files (data the reference's own tests re-run first):

  crates/solver/proptest-regressions/gpu/engine/tests.txt  shrunk SolveConfigs
      of the GPU-engine property tests (crates/solver/src/gpu/engine/tests.rs)
  crates/solver/proptest-regressions/proptests.txt         shrunk river cases
      of crates/solver/src/proptests.rs (board, both ranges, pot / stack)

Only the inputs are kept (boards, 1326-weight ranges, root betting state,
menu, rake, DCFR parameters). Run from the repository root:
    python tests/golden/make_regressions.py [/root/reference]
"""
import json
import os
import re
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "regressions.json")
RANKS, SUITS = "23456789TJQKA", "cdhs"
STREETS = {"Flop": 1, "Turn": 2, "River": 3}


def card(s):
    return RANKS.index(s[0]) * 4 + SUITS.index(s[1])


def cards(txt):
    return [card(x.strip()) for x in txt.split(",") if x.strip()]


def floats(txt):
    return [float(x) for x in txt.split(",") if x.strip()]


def sizes(txt):
    out = []
    for m in re.finditer(r"(Pot|Mult|AllIn)(?:\(([0-9.]+)\))?", txt):
        kind = {"Pot": 0, "Mult": 1, "AllIn": 2}[m.group(1)]
        out.append([kind, float(m.group(2) or 0.0)])
    return out


def engine_case(line):
    c = {"source": "crates/solver/proptest-regressions/gpu/engine/tests.txt", "id": line.split()[1][:12]}
    c["board"] = cards(re.search(r"board: \[([^\]]*)\]", line).group(1))
    rr = re.search(r"ranges: \[\[([^\]]*)\], \[([^\]]*)\]\]", line)
    c["ranges"] = [floats(rr.group(1)), floats(rr.group(2))]
    b = re.search(r"root: BettingState \{ street: (\w+), pot: (\d+), commit: \[(\d+), (\d+)\], stack: \[(\d+), (\d+)\], "
                  r"current_bet: (\d+), last_raise_inc: (\d+), num_raises: (\d+), to_act: (\d+), needs: \[(\w+), (\w+)\]", line)
    c["root"] = {"street": STREETS[b.group(1)], "pot": int(b.group(2)), "commit": [int(b.group(3)), int(b.group(4))],
                 "stack": [int(b.group(5)), int(b.group(6))], "current_bet": int(b.group(7)),
                 "last_raise_inc": int(b.group(8)), "num_raises": int(b.group(9)), "to_act": int(b.group(10)),
                 "needs": [b.group(11) == "true", b.group(12) == "true"]}
    menu = {}
    for st in ("flop", "turn", "river"):
        m = re.search(st + r": StreetSizes \{ bet: \[([^\]]*)\], raise: \[([^\]]*)\] \}", line)
        menu[st] = {"bet": sizes(m.group(1)), "raise": sizes(m.group(2))}
    menu["max_raises"] = int(re.search(r"max_raises: (\d+)", line).group(1))
    for k in ("add_allin_threshold", "force_allin_threshold", "merge_threshold"):
        menu[k] = float(re.search(k + r": ([0-9.]+)", line).group(1))
    c["menu"] = menu
    c["rake_rate"] = float(re.search(r"rake_rate: ([0-9.]+)", line).group(1))
    c["rake_cap"] = float(re.search(r"rake_cap: ([0-9.]+)", line).group(1))
    c["isomorphism"] = re.search(r"isomorphism: (\w+)", line).group(1) == "true"
    d = re.search(r"DcfrParams \{ alpha: ([0-9.]+), beta: Some\(([0-9.]+)\), gamma: ([0-9.]+), restart_average: (\w+), "
                  r"unweighted_average: (\w+)", line)
    c["dcfr"] = {"alpha": float(d.group(1)), "beta": float(d.group(2)), "gamma": float(d.group(3)),
                 "restart_average": d.group(4) == "true", "unweighted_average": d.group(5) == "true"}
    return c


def river_case(line):
    c = {"source": "crates/solver/proptest-regressions/proptests.txt", "id": line.split()[1][:12]}
    c["board"] = cards(re.search(r"board = \[([^\]]*)\]", line).group(1))
    c["ranges"] = [floats(re.search(r"r0 = \[([^\]]*)\]", line).group(1)),
                   floats(re.search(r"r1 = \[([^\]]*)\]", line).group(1))]
    pot = re.search(r"pot = (\d+)", line)
    stack = re.search(r"stack = (\d+)", line)
    c["pot"] = int(pot.group(1)) if pot else None
    c["stack"] = int(stack.group(1)) if stack else None
    return c


def main():
    cases = {"engine": [], "river": []}
    p = os.path.join(REF, "crates/solver/proptest-regressions/gpu/engine/tests.txt")
    for line in open(p):
        if line.startswith("cc "):
            cases["engine"].append(engine_case(line))
    p = os.path.join(REF, "crates/solver/proptest-regressions/proptests.txt")
    for line in open(p):
        if line.startswith("cc "):
            cases["river"].append(river_case(line))
    for c in cases["engine"] + cases["river"]:
        assert all(len(r) == 1326 for r in c["ranges"]), c["id"]
    json.dump(cases, open(OUT, "w"), indent=None, separators=(",", ":"))
    print(f"{len(cases['engine'])} engine and {len(cases['river'])} river cases -> {OUT}")


if __name__ == "__main__":
    main()
