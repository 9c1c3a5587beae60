"""find_overlap_templates on the GPU.

Restates find_overlap_templates of SF/c_cython/sort_cython.pyx:333-462 (SF =
spikesorting_fullpursuit/): for each unit, in descending order of template
energy, the pair of other units (each with at least 4x its spikes) whose
templates, at some pair of shifts in [-n_pre_inds, n_post_inds], sum closest to
it.  The per-pair filters run on the host with the reference's float32
arithmetic; the shift search over every surviving (unit, pair) runs in the
HIP kernel behind fbp_overlap_residuals, which reproduces numpy's FP64
residual sums bit for bit, so the chosen pair and shifts are the reference's.
There is no CPU fallback.
"""
import ctypes

import numpy as np

from .. import _native


def shifted_template(template, shift, n_chans, n_samples_per_chan):
    """Each channel's samples moved by `shift` (zero filled), as get_shifted_template."""
    L = n_samples_per_chan
    t = np.asarray(template).reshape(n_chans, L)
    out = np.zeros((n_chans, L), dtype=np.float64)
    j = np.arange(L)
    src = j - shift
    ok = (src >= 0) & (src < L)
    out[:, ok] = t[:, src[ok]]
    return out.reshape(-1)


def find_overlap_templates(templates, n_samples_per_chan, n_chans, n_pre_inds, n_post_inds, n_template_spikes, *,
                           device=0):
    """List of [test_unit, best_shifted_template (float64), [n1, n2]] as the reference returns."""
    import torch

    T = np.ascontiguousarray(templates, dtype=np.float32)
    U = T.shape[0]
    if U < 3:
        return []
    L, C = int(n_samples_per_chan), int(n_chans)
    n_pre, n_post = int(n_pre_inds), int(n_post_inds)
    nts = np.asarray(n_template_spikes, dtype=np.int64)
    templates_SS = np.sum(T ** 2, axis=1)
    order = np.argsort(templates_SS)[-1::-1]
    triples = []
    for test_unit in order:
        test = T[test_unit]
        cmin = np.argmin(test) // L
        cmax = np.argmax(test) // L
        testmin = np.amin(test[cmin * L:(cmin + 1) * L])
        testmax = np.amax(test[cmax * L:(cmax + 1) * L])
        for n1 in range(U):
            if nts[n1] < 4 * nts[test_unit] or n1 == test_unit:
                continue
            for n2 in range(n1 + 1, U):
                if nts[n2] < 4 * nts[test_unit] or n2 == test_unit:
                    continue
                if np.amin(T[n1, cmin * L:(cmin + 1) * L]) + np.amin(T[n2, cmin * L:(cmin + 1) * L]) > \
                        0.75 * testmin:
                    continue
                if np.amax(T[n1, cmax * L:(cmax + 1) * L]) + np.amax(T[n2, cmax * L:(cmax + 1) * L]) < \
                        0.75 * testmax:
                    continue
                triples.append((int(test_unit), n1, n2))
    if not triples:
        return []
    if not torch.cuda.is_available():
        raise RuntimeError("No GPU available")
    dev = torch.device("cuda", device)
    Td = torch.from_numpy(T).to(dev)
    tr = torch.from_numpy(np.asarray(triples, dtype=np.int32)).to(dev)
    bv = torch.empty(len(triples), dtype=torch.float64, device=dev)
    bq = torch.empty(len(triples), dtype=torch.int32, device=dev)
    vp = ctypes.c_void_p
    rc = _native.load().fbp_overlap_residuals(vp(Td.data_ptr()), U, C, L, n_pre, n_post, vp(tr.data_ptr()),
                                              len(triples), vp(bv.data_ptr()), vp(bq.data_ptr()),
                                              vp(torch.cuda.current_stream(dev).cuda_stream))
    _native.qa_check(rc)
    best_v, best_q = bv.cpu().numpy(), bq.cpu().numpy()
    S = n_pre + n_post + 1
    out = []
    k = 0
    for test_unit in order:
        best = np.inf
        pick = None
        while k < len(triples) and triples[k][0] == test_unit:
            if best_v[k] < best:  # first minimum over the pairs in loop order
                best = best_v[k]
                pick = k
            k += 1
        if pick is not None:
            _, n1, n2 = triples[pick]
            s1 = int(best_q[pick]) // S - n_pre
            s2 = int(best_q[pick]) % S - n_pre
            summed = shifted_template(T[n1], s1, C, L) + shifted_template(T[n2], s2, C, L)
            out.append([int(test_unit), summed, [n1, n2]])
    return out
