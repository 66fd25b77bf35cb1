"""CPU checks of the kernel sources for a hazard no numerical test reliably shows.

A lane move (DPP: from_left / from_right; and the other cross-lane reads: shuffles, ballots, the wavefront
reductions) must run with every lane of the wavefront active. Placed in the
right operand of && or ||, it runs only on the lanes whose left operand decides nothing, and the compiler may
then also stand in the left operand's value for the moved operand (a test of the lane's own level, say), so
the move reads other lanes' registers as that constant or as zero. The result is a wrong neighbour value with
no fault and, where it feeds a rarely taken path (the draining-time limiter of a waiting cell), no visible
effect on the test cases."""
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LANE_MOVE = re.compile(r"\b(from_(left|right)(_i)?|__shfl(_xor|_up|_down)?|__ballot|__any|__all|wave_(sum|max)(_i)?)\s*\(")


def statements(path):
    src = open(path).read()
    src = re.sub(r"//[^\n]*", "", src)
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    # statements end at ';' or at a block brace
    return [s.strip() for s in re.split(r"[;{}]", src) if s.strip()]


def short_circuit_lane_moves(path):
    bad = []
    for s in statements(path):
        for m in LANE_MOVE.finditer(s):
            head = s[:m.start()]
            # a move nested in the argument of another move is evaluated with it; only && / || / ?: that
            # precede the move at the statement's level make it conditional
            if "&&" in head or "||" in head or "?" in head:
                bad.append(s[:160])
                break
    return bad


def test_no_lane_move_in_short_circuit_operands():
    path = os.path.join(ROOT, "floodsim_amd", "csrc", "sw_kernels.h")
    bad = short_circuit_lane_moves(path)
    assert not bad, "lane moves in conditional operands:\n" + "\n".join(bad)


def test_check_finds_a_conditional_lane_move(tmp_path):
    """the check itself: the form it exists to reject, and the form that replaces it"""
    p = tmp_path / "k.h"
    p.write_text("bool w = lv2 == m + 1 && (from_left_i(lv2) == m || lv1 == m);\n"
                 "const int a = from_left_i(lv2), b = from_right_i(lv2);\n"
                 "bool v = (lv2 == m + 1) & ((a == m) | (b == m));\n"
                 "const bool onW = from_left_i(onE ? 1 : 0) != 0;\n"
                 "float q = c ? 0.f : from_right(v);\n")
    bad = short_circuit_lane_moves(str(p))
    assert len(bad) == 2 and bad[0].startswith("bool w") and bad[1].startswith("float q")
