"""Per-template inputs and outputs of the pursuit that live beside it.

Restates the parts of SF/overlap/neuron_separability.py (SF =
spikesorting_fullpursuit/) that the pursuit path reads or writes:
  * thresholds from templates and per-channel noise covariance (:113-195),
  * set_bp_threshold (:319-331),
  * add_n_spikes, called at the end of binary_pursuit (:334-345).
"""
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
        valid = labels[(labels >= 0) & (labels < n_templates)].astype(np.int64)
        counts += np.bincount(valid, minlength=n_templates)[:n_templates]
    separability_metrics["n_spikes"] = counts
    return None
