"""Independent FP64 NumPy restatement of the binary-pursuit pass loop.

TEST INFRASTRUCTURE ONLY.  Written in a different form from fbp_oracle.c so a
misreading shared by both is less likely:
  * the likelihood is a sliding-window matrix product
    (SF/kernels/binary_pursuit.cl:216-252);
  * the overlap value of a pair of window-truncated templates a, b is the drop
    in residual energy 0.5*(|v|^2 - |v - a - b|^2) over the window
    (equal in exact arithmetic to binary_pursuit.cl:585-631, whose per-channel
    "ss > 0" guard only ever skips a zero);
  * window maxima, flags and the pursuit decisions follow
    binary_pursuit.cl:350-441, 657-875, 918-1063 and the loop of
    SF/overlap/binary_pursuit_parallel.py:907-1355 with the same
    phase-separated meaning as the C restatement.
Intended for tiny inputs only (it materialises every overlap item).
"""
import numpy as np
from numpy.lib.stride_tricks import sliding_window_view


def likelihood_rows(V, T, bias, positions):
    """LIK[t, x] for x in positions (0 where x + L >= N)."""
    C, N = V.shape
    U, _, L = T.shape
    out = np.zeros((U, len(positions)))
    pos = np.asarray(positions, dtype=np.int64)
    ok = pos + L < N
    if ok.any():
        for c in range(C):
            win = sliding_window_view(V[c], L)  # [N-L+1, L]
            out[:, ok] += T[:, c, :] @ win[pos[ok]].T
        out[:, ok] += bias[:, None]
    return out


def shifted_truncated(T_u, shift, L):
    """Template placed with onset `shift` samples after the window start,
    truncated to the window [0, L)."""
    C = T_u.shape[0]
    out = np.zeros((C, L))
    lo = max(0, shift)
    hi = min(L, L + shift)
    if hi > lo:
        out[:, lo:hi] = T_u[:, lo - shift:hi - shift]
    return out


def pursuit(V_in, T, bias, thr, n_max, do_recheck=True, return_residual=False):
    V = np.array(V_in, dtype=np.float64)
    T = np.asarray(T, dtype=np.float64)
    bias = np.asarray(bias, dtype=np.float64)
    thr = np.asarray(thr, dtype=np.float64)
    C, N = V.shape
    U, _, L = T.shape
    n = int(n_max)
    S = 2 * n + 1
    nwin = -(-N // L)
    nwf = N // L
    LIK = np.zeros((U, N))
    best_idx = np.zeros(nwin, np.int64)
    best_lab = np.zeros(nwin, np.int64)
    best_lik = np.zeros(nwin)
    obest_idx = np.zeros(nwin, np.int64)
    obest_lab = np.zeros(nwin, np.int64)
    check = np.zeros(nwin, np.uint8)
    orc = np.zeros(nwin, np.uint8)
    W = np.arange(nwin)
    out_idx, out_lab = [], []
    while True:
        pos = np.concatenate([np.arange(i * L, min((i + 1) * L, N)) for i in W]) if len(W) else np.zeros(0, np.int64)
        if len(pos):
            LIK[:, pos] = likelihood_rows(V, T, bias, pos)
        # window maxima (lexicographic: value, then template, then position)
        for i in W:
            sw, ew = i * L, (i + 1) * L
            if ew >= N - L:
                continue
            start, stop = max(0, sw - L), min(N, ew + L)
            blk = LIK[:, start:stop]
            if (blk[:, sw - start:ew - start] > 0).any():
                check[i] = 1
            m = blk.max()
            if m > 0:
                t, x = np.unravel_index(np.argmax(blk), blk.shape)
                best_lab[i], best_idx[i], best_lik[i] = t, start + x, m
            else:
                best_lik[i] = 0.0
            orc[i] = 1 if (best_lik[i] > 0 and sw <= best_idx[i] < ew) else 0
            obest_idx[i], obest_lab[i] = best_idx[i], best_lab[i]
        flagged = np.nonzero(orc)[0]
        if len(flagged) and do_recheck:
            for i in flagged:
                _recheck(V, T, LIK, thr, n, S, L, N, U, i, best_idx, best_lab, best_lik,
                         obest_idx, obest_lab, orc)
            snap = orc.copy()
            for i in flagged:
                if snap[i] == 0:
                    continue
                sw, ew = i * L, (i + 1) * L
                best_idx[i], best_lab[i] = obest_idx[i], obest_lab[i]
                if obest_idx[i] < sw and i > 1:
                    if snap[i - 2] == 1:
                        orc[i] = 0
                        best_idx[i] = 0
                        check[i - 1] = check[i] = check[i + 1] = 2
                    else:
                        check[i - 2] = 2
                if obest_idx[i] >= ew and i + 2 < nwf:
                    check[i + 2] = 2
        added = []
        has = np.zeros(len(W), bool)
        for w, i in enumerate(W):
            sw, ew = i * L, (i + 1) * L
            ml, lab, idx = (best_lik[i], best_lab[i], best_idx[i]) if ew < N - L else (0.0, 0, 0)
            if (ml > thr[lab] or orc[i] == 1) and ((sw <= idx < ew) or orc[i] == 1):
                has[w] = True
                check[i] = 3
                if i > 0:
                    check[i - 1] = 3
                if i + 1 < nwf:
                    check[i + 1] = 3
                added.append((idx, lab))
        snap = check.copy()
        for w, i in enumerate(W):
            sw, ew = i * L, (i + 1) * L
            if not (ew < N - L) or has[w]:
                continue
            own = snap[i]
            if best_idx[i] < sw and own == 1 and i > 0 and snap[i - 1] == 0:
                own = 0
            if best_idx[i] >= ew and own == 1 and i + 1 < nwf and snap[i + 1] == 0:
                own = 0
            if best_lik[i] <= thr[best_lab[i]] and own < 3:
                own = 0
            check[i] = own
        if not added:
            break
        orc[:] = 0
        W = np.nonzero(check)[0]
        check[:] = 0
        for idx, lab in added:
            hi = min(N, idx + L)
            V[:, idx:hi] -= T[lab][:, : hi - idx]
            out_idx.append(idx)
            out_lab.append(lab)
    res = (np.array(out_idx, np.int64), np.array(out_lab, np.int64))
    if return_residual:
        res = res + (V,)
    return res


def _recheck(V, T, LIK, thr, n, S, L, N, U, i, best_idx, best_lab, best_lik, obest_idx,
             obest_lab, orc):
    lab, b = best_lab[i], best_idx[i]
    lb = LIK[lab, b]
    best_val, best_key = None, None
    if not (lb < thr[lab]) and not (b >= N - L):
        vw = V[:, b:b + L]
        e0 = np.sum(vw * vw)
        for s in range(S):
            for f in range(S):
                af, as_ = b + f - n, b + s - n
                if not (0 <= af < N - L) or not (0 <= as_ < N - L):
                    continue
                if abs(f - s) > L:
                    continue
                a = shifted_truncated(T[lab], f - n, L)
                for u in range(U):
                    if LIK[lab, af] <= 0 and LIK[u, as_] <= 0:
                        continue
                    r = vw - a - shifted_truncated(T[u], s - n, L)
                    val = 0.5 * (e0 - np.sum(r * r))
                    if best_val is None or val > best_val:
                        best_val, best_key = val, (u, f, s)
    bl = best_lik[i]
    if best_val is None or not (best_val > bl) or best_val <= 0:
        orc[i] = 0
        return
    u, f, s = best_key
    af, as_ = b + f - n, b + s - n
    A, B = LIK[lab, af], LIK[u, as_]
    if best_val < A or best_val < B or A <= 0 or B <= 0:
        orc[i] = 0
        return
    if A >= B and A > thr[lab]:
        best_lik[i], obest_lab[i], obest_idx[i] = A, lab, af
    elif B > A and B > thr[u]:
        best_lik[i], obest_lab[i], obest_idx[i] = B, u, as_
    elif A > thr[lab]:
        best_lik[i], obest_lab[i], obest_idx[i] = A, lab, af
    else:
        obest_idx[i], obest_lab[i] = best_idx[i], best_lab[i]
        orc[i] = 0
