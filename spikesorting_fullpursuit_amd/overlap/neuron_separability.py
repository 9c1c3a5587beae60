"""Per-template inputs and outputs of the pursuit that live beside it.

Restates the parts of SF/overlap/neuron_separability.py (SF =
spikesorting_fullpursuit/) that the pursuit path reads or writes:
  * compute_separability_metrics (:113-195), with the quadratic forms on the
    GPU (fbp_template_variances), and the same formula on the host
    (template_thresholds) for the synthetic workloads,
  * set_bp_threshold (:319-331),
  * add_n_spikes, called at the end of binary_pursuit (:334-345).
"""
import ctypes

import numpy as np


def template_thresholds(templates, channel_covariance_mats, n_samples_per_chan, sigma_bp_noise,
                        sigma_bp_CI=None):
    """template_SS, neuron_lower_thresholds (before set_bp_threshold) and
    neuron_lower_CI for flattened templates [U][C*L].

    var_n = sum_c t_c^T Sigma_c t_c ; thr = -0.5*SS + sigma_bp_noise*sqrt(var);
    lower_CI = 0.5*SS - sigma_bp_CI*sqrt(var)  (sigma_bp_CI None -> inf).
    """
    T = np.atleast_2d(np.asarray(templates, dtype=np.float64))
    U = T.shape[0]
    L = int(n_samples_per_chan)
    C = T.shape[1] // L
    SS = np.sum(T ** 2, axis=1)
    var = np.zeros(U)
    for n in range(U):
        for c in range(C):
            t = T[n, c * L:(c + 1) * L]
            var[n] += t @ channel_covariance_mats[c] @ t
    ci = np.inf if sigma_bp_CI is None else sigma_bp_CI
    lower_CI = 0.5 * SS - ci * np.sqrt(var)
    thr = -0.5 * SS + sigma_bp_noise * np.sqrt(var)
    return SS, thr, lower_CI


def set_bp_threshold(separability_metrics):
    """neuron_lower_thresholds <- max(thr, 0, neuron_lower_CI), in place."""
    thr = separability_metrics["neuron_lower_thresholds"]
    ci = separability_metrics["neuron_lower_CI"]
    for n in range(ci.shape[0]):
        thr[n] = max(thr[n], 0)
        thr[n] = max(thr[n], ci[n])
    return separability_metrics


def add_n_spikes(separability_metrics, neuron_labels):
    """separability_metrics["n_spikes"][n] = number of spikes labelled n (float64)."""
    n_templates = separability_metrics["templates"].shape[0]
    labels = np.asarray(neuron_labels)
    counts = np.zeros(n_templates)
    if labels.size:
        if np.issubdtype(labels.dtype, np.integer) and labels.min() >= 0 and labels.max() < n_templates:
            counts += np.bincount(labels, minlength=n_templates)[:n_templates]  # every label valid
        else:
            valid = labels[(labels >= 0) & (labels < n_templates)].astype(np.int64)
            counts += np.bincount(valid, minlength=n_templates)[:n_templates]
    separability_metrics["n_spikes"] = counts
    return None


def compute_separability_metrics(templates, channel_covariance_mats, sort_info, *, device=0):
    """The separability_metrics dict of neuron_separability.py:113-195.

    Keys and values as the reference: bp_n_samples_per_chan, templates (stacked),
    template_SS, template_SS_by_chan (zeros, as the reference leaves it),
    channel_covariance_mats, contamination, peak_channel, channel_p_noise,
    neuron_variances (sum_c t_c^T Sigma_c t_c, FP64 on the GPU),
    neuron_lower_thresholds (-SS/2 + sigma_bp_noise * sqrt(var)) and
    neuron_lower_CI (SS/2 - sigma_bp_CI * sqrt(var)).
    """
    import torch
    from scipy.stats import norm

    from .. import _native

    n_chans = sort_info["n_channels"]
    L = sort_info["n_samples_per_chan"]
    T = np.vstack(templates)
    U = T.shape[0]
    m = {"bp_n_samples_per_chan": L, "templates": T}
    m["template_SS"] = np.sum(T ** 2, axis=1)
    m["template_SS_by_chan"] = np.zeros((U, n_chans))
    m["channel_covariance_mats"] = channel_covariance_mats
    m["contamination"] = np.zeros(U)
    m["peak_channel"] = np.zeros(U, dtype=np.int64)
    m["channel_p_noise"] = np.zeros(n_chans)
    for chan in range(n_chans):
        m["channel_p_noise"][chan] = (sort_info["n_samples"] - sort_info["n_threshold_crossings"][chan]) / \
            sort_info["n_samples"]
    if not torch.cuda.is_available():
        raise RuntimeError("No GPU available")
    dev = torch.device("cuda", device)
    Td = torch.from_numpy(np.ascontiguousarray(T, dtype=np.float64)).to(dev)
    covd = torch.from_numpy(np.ascontiguousarray(np.stack([np.asarray(c, np.float64) for c in channel_covariance_mats]))
                            ).to(dev)
    vard = torch.empty(U, dtype=torch.float64, device=dev)
    vp = ctypes.c_void_p
    _native.inputs_check(_native.load().fbp_template_variances(
        vp(Td.data_ptr()), U, n_chans, L, vp(covd.data_ptr()), vp(vard.data_ptr()),
        vp(torch.cuda.current_stream(dev).cuda_stream)))
    var = vard.cpu().numpy()
    expectation = 0.5 * m["template_SS"]
    sd = np.sqrt(var)
    m["neuron_variances"] = var
    m["neuron_lower_CI"] = expectation - sort_info["sigma_bp_CI"] * sd
    m["neuron_lower_thresholds"] = -1 * expectation + sort_info["sigma_bp_noise"] * sd
    m["peak_channel"] = np.argmax(np.abs(T), axis=1) // L
    m["contamination"] = norm.sf(m["neuron_lower_thresholds"], -expectation, sd)
    return m


def check_template_pair(template_1, template_2, chan_covariance_mats, n_chans, template_samples_per_chan, *,
                        device=0):
    """Probability of confusing two aligned templates (neuron_separability.py:59-110):
    norm.cdf(0, E_L(t1) - E_L(t2), sqrt(var)) with E_L(t1) = t1.t1 / 2,
    E_L(t2) = t1.t2 - t2.t2 / 2 and var = sum_c d_c^T Sigma_c d_c for d = t1 - t2
    (the quadratic form in FP64 on the GPU); 1.0 when var <= 0."""
    import torch
    from scipy.stats import norm

    from .. import _native

    t1 = np.asarray(template_1, dtype=np.float64)
    t2 = np.asarray(template_2, dtype=np.float64)
    E_L_t1 = 0.5 * np.dot(t1, t1)
    E_L_t2 = np.dot(t1, t2) - 0.5 * np.dot(t2, t2)
    if not torch.cuda.is_available():
        raise RuntimeError("No GPU available")
    dev = torch.device("cuda", device)
    d = torch.from_numpy(np.ascontiguousarray(t1 - t2)).to(dev)
    covd = torch.from_numpy(np.ascontiguousarray(np.stack([np.asarray(c, np.float64) for c in chan_covariance_mats]))
                            ).to(dev)
    var = torch.empty(1, dtype=torch.float64, device=dev)
    vp = ctypes.c_void_p
    _native.inputs_check(_native.load().fbp_template_variances(
        vp(d.data_ptr()), 1, int(n_chans), int(template_samples_per_chan), vp(covd.data_ptr()), vp(var.data_ptr()),
        vp(torch.cuda.current_stream(dev).cuda_stream)))
    var_diff = float(var.cpu().numpy()[0])
    if var_diff > 0:
        return norm.cdf(0, E_L_t1 - E_L_t2, np.sqrt(var_diff))
    return 1.0
