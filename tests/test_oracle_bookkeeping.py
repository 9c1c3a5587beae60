"""Known answers for the window bookkeeping of the pass loop.

The reference holds no fixtures for binary_pursuit, so every expected value
below is derived by hand from the reference's own lines (SF =
spikesorting_fullpursuit/):

  * compute_template_maximum_likelihood  SF/kernels/binary_pursuit.cl:350-441
  * check_overlap_reassignments         SF/kernels/binary_pursuit.cl:822-875
  * binary_pursuit (add + pruning)      SF/kernels/binary_pursuit.cl:918-1063
  * next window list, flag resets       SF/overlap/binary_pursuit_parallel.py:1217-1290

Two levels:
  * single steps of the pass loop on hand-built state (every window's flags
    and best spike written out);
  * whole runs on tiny noise-free voltages whose likelihoods, overlap values
    and pass-by-pass flags are worked out in the comments: the per-pass trace
    (spikes added, windows listed, windows flagged for the overlap recheck),
    the spikes in output order and the residual.

test_mutants_fail compiles four plausible misreadings of the .cl into copies
of the oracle and requires each to fail the checks that target its rule.
The GPU path is held to the same traces in tests/test_gpu_bookkeeping.py.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

from oracle import fbp_oracle

# --------------------------------------------------------------------------
# Single steps


def check_window_max(lib_=None):
    """compute_template_maximum_likelihood, launched for t = 0..U-1.

    U = 2, L = 4, N = 40 (10 windows; windows with (id+1)*L >= N-L, ids 8 and 9,
    return at once, .cl:380-383).  The scan of window id covers
    [id*L - L, (id+2)*L) (.cl:366-379); best starts at 0 with label and index
    NOT reset (.cl:386-393); strict > in template-major, index-minor order."""
    L, N, U = 4, 40, 2
    LIK = np.full((U, N), -1.0, np.float32)
    LIK[0, 9] = 3.0    # tie across templates at x = 9: template 0 keeps it
    LIK[1, 9] = 3.0
    LIK[1, 13] = 1.0   # positive in window 3, below the best at 9
    LIK[0, 23] = 4.0   # template 0 first, then template 1 beats it at 21
    LIK[1, 21] = 5.0
    LIK[1, 22] = 5.0   # tie across positions: the first (21) is kept
    LIK[0, 29] = 0.0   # zero is not positive
    LIK[0, 33] = 9.0   # inside window 8, which itself returns at once
    st = fbp_oracle.new_state(10)
    st["best_lab"][0], st["best_idx"][0] = 1, 77      # stale values survive
    st["best_lik"][8] = 0.5                           # window 8 untouched
    st["check"][1], st["check"][6] = 2, 3             # unlisted windows untouched
    win = [0, 2, 3, 4, 5, 7, 8]
    fbp_oracle.window_max_step(LIK, L, win, st, lib_=lib_)
    # (best_lik, best_lab, best_idx, check, orc) per window
    want = {
        0: (0.0, 1, 77, 0, 0),  # nothing positive in [0, 8)
        1: (0.0, 0, 0, 2, 0),   # not listed
        2: (3.0, 0, 9, 1, 1),   # best in own window [8, 12)
        3: (3.0, 0, 9, 1, 0),   # own positive at 13, best in window 2
        4: (5.0, 1, 21, 0, 0),  # 4.0 at 23 (t=0) then 5.0 at 21 (t=1); none in [16, 20)
        5: (5.0, 1, 21, 1, 1),
        6: (0.0, 0, 0, 3, 0),   # not listed
        7: (9.0, 0, 33, 0, 0),  # best inside the returning window 8; 0 at 29 is no positive
        8: (0.5, 0, 0, 0, 0),   # (8+1)*4 >= 40-4: returns at once
        9: (0.0, 0, 0, 0, 0),
    }
    for w, (lik, lab, idx, chk, orc) in want.items():
        got = (float(st["best_lik"][w]), int(st["best_lab"][w]), int(st["best_idx"][w]),
               int(st["check"][w]), int(st["orc"][w]))
        assert got == (lik, lab, idx, chk, orc), (w, got)
        if w not in (1, 6, 8, 9):  # overlap_best_* copies the best (.cl:435-439)
            assert (int(st["obest_idx"][w]), int(st["obest_lab"][w])) == (idx, lab), w


def check_reassign(lib_=None):
    """check_overlap_reassignments (.cl:822-875), overlap_recheck read as it
    stood before the kernel.  N = 395, L = 10: 40 windows and
    voltage_length / template_length = 39 for the right-move bound (.cl:870)."""
    N, L = 395, 10
    st = fbp_oracle.new_state(40)
    flist = [1, 3, 5, 8, 9, 13, 17, 37]
    orc = {1: 1, 3: 1, 5: 1, 8: 1, 9: 1, 13: 0, 17: 1, 37: 1}
    obest = {1: (8, 1), 3: (28, 0), 5: (47, 1), 8: (83, 1), 9: (85, 0), 13: (125, 1), 17: (181, 0),
             37: (381, 1)}
    for w in flist:
        st["orc"][w] = orc[w]
        st["check"][w] = 1
        st["best_idx"][w], st["best_lab"][w] = w * L + 2, 0
        st["obest_idx"][w], st["obest_lab"][w] = obest[w]
    fbp_oracle.reassign_step(N, L, flist, st, lib_=lib_)
    want_check = {
        1: 1,              # left of window 1 but id > 1 fails (.cl:850): nothing
        2: 2, 3: 2, 4: 2,  # window 3 moves left, overlap_recheck[1] == 1: gives up (.cl:852-862)
        5: 2, 6: 2,        # window 5 moves left, overlap_recheck[3] was 1 before the kernel: gives up
        7: 2,              # window 9 moves left, overlap_recheck[7] == 0: flags id-2 (.cl:864-866)
        8: 1, 9: 1, 13: 1, 17: 1,
        19: 2,             # window 17 moves right: flags id+2 (.cl:870-873)
        37: 1,             # window 37 moves right but 37 + 2 < 39 fails
    }
    got_check = {w: int(c) for w, c in enumerate(st["check"]) if c}
    assert got_check == want_check
    want_orc = {1: 1, 3: 0, 5: 0, 8: 1, 9: 1, 13: 0, 17: 1, 37: 1}
    assert {w: int(st["orc"][w]) for w in flist} == want_orc
    want_best = {1: (8, 1), 3: (0, 0), 5: (0, 1), 8: (83, 1), 9: (85, 0), 13: (132, 0), 17: (181, 0),
                 37: (381, 1)}
    assert {w: (int(st["best_idx"][w]), int(st["best_lab"][w])) for w in flist} == want_best


def check_pursuit(lib_=None):
    """binary_pursuit (.cl:918-1063): every "=3" write before any pruning.
    N = 400, L = 10: 40 windows, valid while (id+1)*L < N-L (ids 0..37);
    thresholds 1.0 (unit 0) and 5.0 (unit 1)."""
    N, L = 400, 10
    thr = np.float32([1.0, 5.0])
    st = fbp_oracle.new_state(40)
    # window: (best_lik, best_lab, best_idx, check, overlap_recheck)
    pre = {
        0: (3.0, 0, 4, 1, 0),      # added; id 0 has no left neighbour
        2: (0.5, 0, 25, 1, 0),     # below threshold: pruned to 0 (.cl:1039-1044)
        5: (4.0, 0, 45, 1, 0),     # best in window 4, flag[4] == 0: pruned (.cl:1024-1030)
        8: (4.0, 0, 75, 1, 0),     # best in window 7, flag[7] == 2: kept
        11: (4.0, 0, 123, 1, 0),   # best in window 12, flag[12] == 0: pruned (.cl:1032-1038)
        14: (0.2, 1, 138, 1, 1),   # overlap recheck: added below threshold, left of its window
        17: (0.5, 1, 175, 2, 0),   # below unit 1's threshold, flag 2 < 3: pruned
        20: (6.0, 1, 203, 1, 0),   # added
        23: (4.0, 0, 245, 1, 0),   # best in window 24, which window 25's spike flags 3: kept
        25: (2.0, 0, 252, 1, 0),   # added
        30: (1.0, 0, 305, 1, 0),   # 1.0 > 1.0 fails; pruned by the threshold rule
        32: (4.0, 0, 315, 2, 0),   # best in window 31 but flag 2, not 1: kept
        38: (9.0, 0, 385, 1, 0),   # (38+1)*10 >= 400-10: treated as empty, never pruned
    }
    for w, (lik, lab, idx, chk, orc) in pre.items():
        st["best_lik"][w], st["best_lab"][w], st["best_idx"][w] = lik, lab, idx
        st["check"][w], st["orc"][w] = chk, orc
    st["check"][7] = 2  # not listed
    win = sorted(pre)
    idx, lab = fbp_oracle.pursuit_step(N, L, thr, win, st, lib_=lib_)
    assert list(zip(idx.tolist(), lab.tolist())) == [(4, 0), (138, 1), (203, 1), (252, 0)]
    want_check = {0: 3, 1: 3, 7: 2, 8: 1, 13: 3, 14: 3, 15: 3, 19: 3, 20: 3, 21: 3, 23: 1, 24: 3, 25: 3,
                  26: 3, 32: 2, 38: 1}
    got_check = {w: int(c) for w, c in enumerate(st["check"]) if c}
    assert got_check == want_check


# --------------------------------------------------------------------------
# Whole runs


def case_prune_rules():
    """Two channels, L = 4, N = 128 (32 windows; ids 30, 31 return at once),
    no shift (n_max = 0), recheck on.  Unit 0 = 2 at onset on channel 0,
    unit 1 = 2 at onset on channel 1; bias -2 each, thresholds 1 and 10, so
    LIK_u[x] = 2 V_u[x] - 2 (exact) and -2 wherever the voltage is 0.

      channel 0: V[9] = 2 (LIK 2), V[13] = 1.75 (1.5), V[42] = 1.25 (0.5),
                 V[89] = 1.75 (1.5), V[93] = 1.25 (0.5), V[117] = 1.25 (0.5),
                 V[121] = 2 (2)
      channel 1: V[85] = 3 (LIK_1 4, below unit 1's threshold 10)

    Pass 1, all 32 windows.  Flagged (best inside own window): 2 (2 at 9),
    10 (0.5 at 42), 21 (4 at 85, unit 1 after unit 0's 1.5 at 89) -> 3.
    Recheck: window 2 evaluates u = 0 (2+2 at 9: 8 - 8 = 0) and u = 1
    (2*2 - 2 + 0 - 2 = 0); 0 > 2 fails, flag cleared.  Windows 10 and 21 are
    below threshold: every item skipped (.cl:518-522), flag cleared.
    Window 2 adds (9, 0) and flags 1-3 with 3.  Pruning, flags after the
    3-writes: 10, 21, 22 below threshold -> 0; 23 (best 1.5 at 89 in window
    22, own 0.5 at 93) keeps 1 because window 22 was flagged 1; 29 (best 2 at
    121 in window 30, own 0.5 at 117) -> 0 because window 30 returned at
    once and holds 0 (.cl:1032-1038).  Next list: 1, 2, 3, 23.
    Pass 2 (V[9] = 0): flagged 3 (1.5 at 13); its recheck values -1 and -0.5
    do not beat 1.5.  Window 3 adds (13, 0).  Window 23: best 1.5 at 89 in
    window 22, which is not listed (flag 0): pruned (.cl:1024-1030).  Window
    1: best 0 -> 0.  Next list: 2, 3, 4.
    Pass 3 (V[13] = -0.25): nothing positive, nothing added: the loop ends."""
    C, N, U, L = 2, 128, 2, 4
    T = np.zeros((U, C, L), np.float32)
    T[0, 0, 0] = 2.0
    T[1, 1, 0] = 2.0
    V = np.zeros((C, N), np.float32)
    for x, v in ((9, 2.0), (13, 1.75), (42, 1.25), (89, 1.75), (93, 1.25), (117, 1.25), (121, 2.0)):
        V[0, x] = v
    V[1, 85] = 3.0
    bias = np.float32([-2.0, -2.0])
    thr = np.float32([1.0, 10.0])
    R = V.copy()
    R[0, 9] = 0.0
    R[0, 13] = -0.25
    return dict(V=V, T=T, bias=bias, thr=thr, n=0, trace=[(1, 32, 3), (1, 4, 1), (0, 3, 0)],
                spikes=[(9, 0), (13, 0)], residual=R)


def _merged(a1, a2):
    """Unit [1, 2, 2, 1] at p scaled by a1 plus unit at p + 2 scaled by a2."""
    return np.array([a1, 2 * a1, 2 * a1 + a2, a1 + 2 * a2, 2 * a2, a2], np.float32)


def case_left_moves():
    """One channel, one unit T = [1, 2, 2, 1] (bias -5, threshold 1), L = 4,
    N = 48 (12 windows; ids 10, 11 return at once), n_max = 1 (S = 3).

    Two merged pairs 1.25*T(p) + T(p+2), at p = 15 and p = 23:
    V[p..p+5] = [1.25, 2.5, 3.5, 3.25, 2, 1].  LIK[x] = sum T*V[x..x+3] - 5:
    p-2: 0, p-1: 6, p: 11.5, p+1: 13, p+2: 10, p+3: 4.25, p+4: -1.
    The merged peak p+1 starts windows 4 (p = 15) and 6 (p = 23).

    Overlap values over the window [b, b+4), b = p+1, with the unit placed at
    b-1, b, b+1 (truncated rows [2,2,1,0], [1,2,2,1], [0,1,2,2]) and
    V[b..b+3] = [2.5, 3.5, 3.25, 2]: (-1,-1) 12.5, (-1,0) 15.75, (-1,+1)
    16.25, (0,0) 16, (0,+1) 14.5, (+1,+1) 10.  The maximum 16.25 first occurs
    at item k = f + 3 s = 2 (fixed at b+1, other at b-1); 16.25 > 13.  Cascade
    (.cl:733-814): A = LIK[b+1] = 10, B = LIK[b-1] = 11.5 > A and > 1: the
    spike moves to b-1 = p, left of its window, best_lik 11.5.

    Pass 1 (12 windows): flagged 4 and 6 (2).  Reassignment: window 4 moves
    to 15; overlap_recheck[2] == 0 so window 2 is flagged 2 (.cl:864-866).
    Window 6 moves to 23; overlap_recheck[4] == 1 so it gives up: flag
    cleared, best index 0, windows 5-7 flagged 2 (.cl:852-862).  Window 4
    adds (15, 0) (overlap flag) and flags 3-5 with 3; window 6 is not added
    (index 0 outside, flag cleared) and keeps 2.  Next list 2..7 (6).
    Pass 2 (pair at 15 reduced to 0.25*T(15) + T(17): LIK 15: 1.5, 16: 5,
    17: 6, 18: 3.25): only window 6 is flagged; it moves left again, now with
    overlap_recheck[4] == 0: flags window 4 with 2 and adds (23, 0).  Window
    3 (best 6 at 17 in window 4, own 1.5 at 15) keeps 1: window 4 holds 2.
    Window 2 (best 1.5 at 15) is pruned.  Next list 3..7 (5).
    Pass 3: both reduced pairs peak at 6 (17 and 25) in windows 4 and 6
    (flagged 2); their best overlap value 4.75 < 6 is rejected; both are
    added, (17, 0) then (25, 0).  Next list 3..7.
    Pass 4: 0.25*T at 15 and 23 remain, every likelihood is negative."""
    N, L = 48, 4
    T = np.float32([[[1, 2, 2, 1]]])
    V = np.zeros((1, N), np.float32)
    V[0, 15:21] = _merged(1.25, 1.0)
    V[0, 23:29] = _merged(1.25, 1.0)
    R = np.zeros((1, N), np.float32)
    R[0, 15:19] = [0.25, 0.5, 0.5, 0.25]
    R[0, 23:27] = [0.25, 0.5, 0.5, 0.25]
    return dict(V=V, T=T, bias=np.float32([-5.0]), thr=np.float32([1.0]), n=1,
                trace=[(1, 12, 2), (1, 6, 1), (2, 5, 2), (0, 5, 0)],
                spikes=[(15, 0), (23, 0), (17, 0), (25, 0)], residual=R)


def case_right_move():
    """Same unit and shift range, N = 32 (8 windows; ids 6, 7 return at once).
    One merged pair T(10) + T(12): V[10..15] = [1, 2, 3, 3, 2, 1]; LIK 9: 4,
    10: 9, 11: 11, 12: 9, 13: 4.  The peak 11 is the last sample of window 2.
    Overlap values at b = 11 over V[11..14] = [2, 3, 3, 2]: (-1,-1) 8, (-1,0)
    11.5, (-1,+1) 13, (0,0) 12, (0,+1) 11.5, (+1,+1) 8; first maximum k = 2
    (fixed at 12, other at 10); 13 > 11.  Cascade: A = LIK[12] = 9 >= B =
    LIK[10] = 9 and > 1: the spike moves to 12, right of window 2.

    Pass 1 (8 windows, 1 flagged): window 2 flags window 4 with 2
    (.cl:870-873) and adds (12, 0), flagging 1-3 with 3.  Window 4 (best 9 at
    12, flag 2) is kept.  Next list 1..4.
    Pass 2 (T(10) left): window 2 flagged (5 at 10); overlap values at most 3,
    rejected; (10, 0) added.  Window 4 (nothing positive in its scan) is
    pruned.  Next list 1..3.  Pass 3: the residual is 0, nothing added."""
    N, L = 32, 4
    T = np.float32([[[1, 2, 2, 1]]])
    V = np.zeros((1, N), np.float32)
    V[0, 10:16] = _merged(1.0, 1.0)
    return dict(V=V, T=T, bias=np.float32([-5.0]), thr=np.float32([1.0]), n=1,
                trace=[(1, 8, 1), (1, 4, 1), (0, 3, 0)], spikes=[(12, 0), (10, 0)],
                residual=np.zeros((1, N), np.float32))


CASES = {"prune_rules": case_prune_rules, "left_moves": case_left_moves, "right_move": case_right_move}


def check_case(name, lib_=None):
    c = CASES[name]()
    idx, lab, res, log = fbp_oracle.pursuit(c["V"], c["T"], c["bias"], c["thr"], c["n"], do_recheck=True,
                                            return_residual=True, return_log=True, lib_=lib_)
    assert [tuple(r) for r in log[:, :3].tolist()] == c["trace"]
    assert list(zip(idx.tolist(), lab.tolist())) == c["spikes"]
    np.testing.assert_array_equal(res, c["residual"])


# --------------------------------------------------------------------------
# Tests


def test_window_max_step_known_answer():
    check_window_max()


def test_reassign_step_known_answer():
    check_reassign()


def test_pursuit_step_known_answer():
    check_pursuit()


@pytest.mark.parametrize("name", sorted(CASES))
def test_pass_trace_known_answer(name):
    check_case(name)


def test_case_likelihoods_as_stated():
    # The likelihood values the derivations above rest on, from the closed
    # form bias + sum_c correlate(V_c, T_c) (binary_pursuit.cl:216-252).
    c = case_left_moves()
    lik = np.correlate(c["V"][0].astype(np.float64), c["T"][0, 0].astype(np.float64), "valid") - 5.0
    for p in (15, 23):
        assert lik[p - 2:p + 5].tolist() == [0.0, 6.0, 11.5, 13.0, 10.0, 4.25, -1.0]
    c = case_right_move()
    lik = np.correlate(c["V"][0].astype(np.float64), c["T"][0, 0].astype(np.float64), "valid") - 5.0
    assert lik[9:14].tolist() == [4.0, 9.0, 11.0, 9.0, 4.0]
    # overlap values as the residual-energy drop 0.5 (|v|^2 - |v - a - b|^2)
    v = np.array([2.5, 3.5, 3.25, 2.0])
    rows = {-1: np.array([2, 2, 1, 0.0]), 0: np.array([1, 2, 2, 1.0]), 1: np.array([0, 1, 2, 2.0])}
    drop = {(a, b): 0.5 * (v @ v - np.sum((v - rows[a] - rows[b]) ** 2)) for a in rows for b in rows}
    assert drop[(1, -1)] == drop[(-1, 1)] == 16.25 == max(drop.values())
    assert sorted(set(drop.values()))[-2] == 16.0


# Misreadings of the reference that a shared oracle could carry: each is
# planted in a copy of the oracle and must fail the checks aimed at its rule.
MUTANTS = {
    # A: drop "best <= threshold and flag < 3 -> 0" (.cl:1039-1044)
    "A": (["        if (ml <= thr[lab] && own < 3) own = 0;\n"], [""],
          ["pursuit", "prune_rules"]),
    # B: test overlap_recheck[id-1] instead of [id-2] on a left move (.cl:852)
    "B": (["            if (orc_snap[id - 2] == 1) {\n"], ["            if (orc_snap[id - 1] == 1) {\n"],
          ["reassign", "left_moves"]),
    # C: drop both neighbour pruning rules (.cl:1024-1038)
    "C": (["        if (idx < sw && own == 1 && id > 0 && check_snap[id - 1] == 0) own = 0;\n",
           "        if (idx >= ew && own == 1 && id + 1 < nwf && check_snap[id + 1] == 0) own = 0;\n"],
          ["", ""], ["pursuit", "prune_rules"]),
    # D: flag id+1 instead of id+2 on a right move (.cl:872)
    "D": (["        if ((int64_t)st->obest_idx[id] >= ew && id + 2 < nwf) st->check[id + 2] = 2;\n"],
          ["        if ((int64_t)st->obest_idx[id] >= ew && id + 2 < nwf) st->check[id + 1] = 2;\n"],
          ["reassign", "right_move"]),
}

CHECKS = {
    "window_max": check_window_max,
    "reassign": check_reassign,
    "pursuit": check_pursuit,
    "prune_rules": lambda lib_: check_case("prune_rules", lib_),
    "left_moves": lambda lib_: check_case("left_moves", lib_),
    "right_move": lambda lib_: check_case("right_move", lib_),
}


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc to build the mutant oracles")
@pytest.mark.parametrize("mutant", sorted(MUTANTS))
def test_mutants_fail(mutant, tmp_path):
    src = open(os.path.join(os.path.dirname(fbp_oracle.__file__), "fbp_oracle.c")).read()
    olds, news, must_fail = MUTANTS[mutant]
    for old, new in zip(olds, news):
        assert src.count(old) == 1, old
        src = src.replace(old, new)
    c_path = tmp_path / f"mutant_{mutant}.c"
    so_path = tmp_path / f"libmutant_{mutant}.so"
    c_path.write_text(src)
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-fopenmp", "-fPIC", "-shared", "-o", str(so_path),
                    str(c_path), "-lm"], check=True)
    lib_ = fbp_oracle.bind(str(so_path))
    failed = set()
    for name, fn in CHECKS.items():
        try:
            fn(lib_)
        except AssertionError:
            failed.add(name)
    assert set(must_fail) <= failed, (mutant, sorted(failed))
