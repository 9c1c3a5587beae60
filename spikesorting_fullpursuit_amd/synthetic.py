"""Seeded synthetic recordings for the binary-pursuit workloads.

Every workload is float32 voltage = band-limited Gaussian noise (sd 1) plus
difference-of-Gaussians templates placed at Poisson spike times:

  small      tetrode, 4 ch, 40 kHz, 30 s, 5 units, 64-sample templates
  probe16    16-ch line (50 um pitch), 40 kHz, 300 s, 40 units on 5 adjacent
             channels, noise with 0.3 cross-channel correlation
  probe64    64-ch line (25 um pitch), 30 kHz, 600 s, 120 units, 48-sample
             templates on 7-channel neighbourhoods, rates 1-50 Hz
  probe32    32-ch line, 30 kHz, 120 s, 60 units, 64-sample templates

The per-template inputs of the pursuit (template_SS, neuron_lower_thresholds)
are computed as SF/overlap/neuron_separability.py:113-195 and :319-331 do:
  template_SS = sum(t^2)
  var         = sum_c t_c^T Sigma_c t_c     (Sigma_c: L x L noise covariance)
  thr         = max(-0.5*SS + sigma_bp_noise*sqrt(var), 0, 0.5*SS - sigma_bp_CI*sqrt(var))
with sigma_bp_noise = 3.719 and sigma_bp_CI = inf (SF/parallel/spikesorting_parallel.py:76-77).
"""
import functools
from dataclasses import dataclass, field

import numpy as np
from scipy import signal

SIGMA_BP_NOISE = 3.719
SIGMA_BP_CI = np.inf


@dataclass
class Workload:
    name: str
    n_chans: int
    fs: float
    duration_s: float
    n_units: int
    L: int
    chan_xy: np.ndarray          # [C, 2] um
    neighbourhood: int           # channels per template (0 = all)
    rate_range: tuple
    noise_corr: float = 0.0
    seed: int = 20261200
    extra: dict = field(default_factory=dict)

    @property
    def n_samples(self):
        return int(round(self.fs * self.duration_s))


def workload(name, duration_s=None, seed=None):
    """Return the Workload description for `name` (optionally shortened)."""
    if name == "small":
        xy = np.array([[0.0, 0.0], [20.0, 0.0], [0.0, 20.0], [20.0, 20.0]])
        w = Workload("small", 4, 40000.0, 30.0, 5, 64, xy, 0, (2.0, 20.0), 0.0, 20261201)
    elif name == "probe16":
        xy = np.stack([np.zeros(16), 50.0 * np.arange(16)], axis=1)
        w = Workload("probe16", 16, 40000.0, 300.0, 40, 64, xy, 5, (2.0, 80.0), 0.3, 20261202)
    elif name == "probe64":
        xy = np.stack([np.zeros(64), 25.0 * np.arange(64)], axis=1)
        w = Workload("probe64", 64, 30000.0, 600.0, 120, 48, xy, 7, (1.0, 50.0), 0.0, 20261203)
    elif name == "probe32":
        xy = np.stack([np.zeros(32), 25.0 * np.arange(32)], axis=1)
        w = Workload("probe32", 32, 30000.0, 120.0, 60, 64, xy, 7, (1.0, 50.0), 0.0, 20261205)
    else:
        raise ValueError(f"unknown workload {name}")
    if duration_s is not None:
        w.duration_s = float(duration_s)
    if seed is not None:
        w.seed = int(seed)
    return w


def dog_waveform(L, fs, rng):
    """Difference of Gaussians (sigmas 0.15 ms and 0.4 ms), trough-dominated,
    normalised to unit absolute peak."""
    t = (np.arange(L) - L * 0.45) / fs
    s1, s2 = 0.15e-3, 0.4e-3
    shift = rng.uniform(-0.05e-3, 0.05e-3)
    g1 = np.exp(-0.5 * ((t - shift) / s1) ** 2)
    g2 = np.exp(-0.5 * ((t - shift - 0.2e-3) / s2) ** 2)
    w = -g1 + 0.45 * g2
    return w / np.max(np.abs(w))


def make_templates(w, rng):
    """[U][C][L] float64 templates, zero outside each unit's neighbourhood."""
    C, U, L = w.n_chans, w.n_units, w.L
    T = np.zeros((U, C, L))
    for u in range(U):
        wave = dog_waveform(L, w.fs, rng)
        amp = rng.uniform(4.0, 12.0)
        if w.neighbourhood == 0:
            centre = w.chan_xy.mean(axis=0)
            ang = rng.uniform(0, 2 * np.pi)
            rad = rng.uniform(0, 30.0)
            pos = centre + rad * np.array([np.cos(ang), np.sin(ang)])
            chans = np.arange(C)
        else:
            h = w.neighbourhood // 2
            c0 = int(rng.integers(h, C - h))
            pos = w.chan_xy[c0] + np.array([rng.uniform(10.0, 30.0),
                                            rng.uniform(-0.5, 0.5) * (w.chan_xy[1, 1] - w.chan_xy[0, 1])])
            chans = np.arange(c0 - h, c0 - h + w.neighbourhood)
        d = np.sqrt(np.sum((w.chan_xy[chans] - pos) ** 2, axis=1))
        gain = np.exp(-d / 40.0)
        gain = gain / gain.max()
        T[u, chans, :] = amp * gain[:, None] * wave[None, :]
    return T


def spike_trains(w, rng, n_samples):
    """Poisson trains with a 1.5 ms refractory period; returns (onsets, labels, amp)."""
    refr = 1.5e-3
    onsets, labels = [], []
    dur = n_samples / w.fs
    for u in range(w.n_units):
        rate = rng.uniform(*w.rate_range)
        mean_isi = 1.0 / rate
        n_exp = int(dur * rate * 1.3) + 20
        isi = refr + rng.exponential(max(mean_isi - refr, 1e-6), size=n_exp)
        t = np.cumsum(isi)
        t = t[t < dur]
        idx = np.round(t * w.fs).astype(np.int64)
        idx = idx[(idx >= 0) & (idx + w.L < n_samples)]
        onsets.append(idx)
        labels.append(np.full(idx.shape, u, np.int64))
    onsets = np.concatenate(onsets)
    labels = np.concatenate(labels)
    order = np.lexsort((labels, onsets))
    onsets, labels = onsets[order], labels[order]
    amp = rng.normal(1.0, 0.05, size=onsets.shape)
    return onsets, labels, amp


def lowpass_sos(fs):
    return signal.butter(4, 8000.0, btype="low", fs=fs, output="sos")


def make_noise(w, rng, n_samples):
    """Band-limited noise, unit sd per channel, optional shared component."""
    sos = lowpass_sos(w.fs)

    def one():
        x = signal.sosfiltfilt(sos, rng.standard_normal(n_samples))
        return x / np.std(x)

    shared = one() if w.noise_corr > 0 else None
    out = np.empty((w.n_chans, n_samples), np.float32)
    for c in range(w.n_chans):
        x = one()
        if shared is not None:
            x = np.sqrt(1 - w.noise_corr) * x + np.sqrt(w.noise_corr) * shared
        out[c] = x
    return out


def noise_autocov(fs, L, n=1 << 18, seed=0):
    """Autocovariance (lags 0..L-1) of the unit-sd band-limited noise, from the
    squared magnitude response of the zero-phase filter (closed form)."""
    sos = lowpass_sos(fs)
    wgrid, h = signal.sosfreqz(sos, worN=n, fs=fs)
    psd = np.abs(h) ** 4
    ac = np.fft.irfft(psd)
    ac = ac / ac[0]
    return ac[:L]


def separability_inputs(T, fs, sigma_bp_noise=SIGMA_BP_NOISE, sigma_bp_CI=SIGMA_BP_CI):
    """template_SS, template_SS_by_chan and neuron_lower_thresholds for [U][C][L] T."""
    U, C, L = T.shape
    ac = noise_autocov(fs, L)
    idx = np.abs(np.arange(L)[:, None] - np.arange(L)[None, :])
    Sigma = ac[idx]
    SS = np.sum(T.reshape(U, -1) ** 2, axis=1)
    SS_by_chan = np.sum(T ** 2, axis=2)
    var = np.einsum("uci,ij,ucj->u", T, Sigma, T)
    thr = -0.5 * SS + sigma_bp_noise * np.sqrt(var)
    lower_ci = 0.5 * SS - sigma_bp_CI * np.sqrt(var)
    thr = np.maximum(np.maximum(thr, 0.0), lower_ci)
    return SS, SS_by_chan, thr


def clip_width_for(L, fs):
    """clip_width (s) whose sample window has exactly L samples."""
    pre = L // 2
    post = L - pre - 1
    return [-pre / fs, post / fs], pre


BATCH_CHANNELS = (1, 4, 8, 16, 32)


def batch_workloads(n_segments=256, seed=20261204, duration_range=(5.0, 120.0), unit_range=(4, 60),
                    channels=BATCH_CHANNELS, fs=30000.0, L=64):
    """The batched workload: independent segments with lengths drawn uniformly
    from duration_range (s), channel counts from `channels` and unit counts
    from unit_range (inclusive), 64-sample templates at 30 kHz.  Channels sit
    on a line at 25 um pitch; units span all channels of a segment with at most
    4 channels and 5 adjacent channels otherwise; rates 1-30 Hz."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n_segments):
        dur = float(rng.uniform(*duration_range))
        C = int(rng.choice(channels))
        U = int(rng.integers(unit_range[0], unit_range[1] + 1))
        xy = np.stack([np.zeros(C), 25.0 * np.arange(C)], axis=1)
        nb = 0 if C <= 4 else 5
        out.append(Workload(f"segment{i}", C, fs, dur, U, L, xy, nb, (1.0, 30.0), 0.0, seed + 1 + i))
    return out


def generate(name, duration_s=None, seed=None):
    """Build a workload.  Returns a dict with voltage [C][N] float32, templates
    [U][C][L] float64, truth (onsets, labels), sort_info and separability_metrics
    in the form SF/overlap/binary_pursuit_parallel.py:175-183 consumes."""
    return generate_workload(workload(name, duration_s, seed))


def generate_workload(w):
    """generate() for a Workload description."""
    rng = np.random.default_rng(w.seed)
    N = w.n_samples
    T = make_templates(w, rng)
    onsets, labels, amp = spike_trains(w, rng, N)
    V = make_noise(w, rng, N)
    add_spikes(V, T, onsets, labels, amp)
    return package(w, V, T, onsets, labels)


def add_spikes(V, T, onsets, labels, amp):
    U, C, L = T.shape
    Tf = T.astype(np.float32)
    for u in range(U):
        sel = labels == u
        if not sel.any():
            continue
        on = onsets[sel]
        a = amp[sel].astype(np.float32)
        rows = on[:, None] + np.arange(L)[None, :]
        for c in range(C):
            if not np.any(Tf[u, c]):
                continue
            np.add.at(V[c], rows, a[:, None] * Tf[u, c][None, :])


def package(w, V, T, onsets, labels):
    U, C, L = T.shape
    SS, SS_by_chan, thr = separability_inputs(T, w.fs)
    clip_width, pre = clip_width_for(L, w.fs)
    sort_info = {
        "sampling_rate": w.fs,
        "clip_width": clip_width,
        "n_samples_per_chan": L,
        "n_channels": C,
        "n_samples": V.shape[1],
        "do_overlap_recheck": True,
        "get_adjusted_clips": False,
        "max_gpu_memory": None,
    }
    metrics = {
        "templates": T.reshape(U, C * L).copy(),
        "template_SS": SS,
        "template_SS_by_chan": SS_by_chan,
        "neuron_lower_thresholds": thr,
    }
    return {
        "workload": w,
        "voltage": V,
        "templates": T,
        "truth_onsets": onsets,
        "truth_labels": labels,
        "sort_info": sort_info,
        "separability_metrics": metrics,
        "clip_init_samples": pre,
        "n_max_shift_inds": L - 1,
    }


def kernel_inputs(data):
    """(T [U][C][L] float32, bias float32, thr float32) exactly as the drop-in
    derives them (SF/overlap/binary_pursuit_parallel.py:239-264)."""
    m = data["separability_metrics"]
    C = data["voltage"].shape[0]
    tm = np.float32(m["templates"])
    U = tm.shape[0]
    T = tm.reshape(U, C, -1)
    bias = np.float32(-0.5 * m["template_SS"])
    thr = np.float32(m["neuron_lower_thresholds"])
    return np.ascontiguousarray(T), bias, thr


def generate_on_device(name, duration_s=None, seed=None, device="cuda"):
    """Same workload as generate() built on the GPU with torch (for the large
    benchmark recordings): templates, spike trains and thresholds come from the
    same seeded NumPy code; the noise is white Gaussian shaped in the frequency
    domain by the zero-phase filter's squared magnitude response and rescaled
    to sd 1.  Returns the dict of generate() with `voltage` a float32 device
    tensor [C][N] (the realisation differs from generate()'s)."""
    return generate_workload_on_device(workload(name, duration_s, seed), device)


@functools.lru_cache(maxsize=8)
def _power_gain(fs, nfft):
    """|H(f)|^2 of the zero-phase low-pass at the rfft frequencies of length nfft."""
    freqs = np.fft.rfftfreq(nfft, d=1.0 / fs)
    _, h = signal.sosfreqz(lowpass_sos(fs), worN=freqs, fs=fs)
    return (np.abs(h) ** 2).astype(np.float32)


def generate_workload_on_device(w, device="cuda", pow2_fft=False):
    """generate_on_device() for a Workload description.  pow2_fft shapes the
    noise on the next power-of-two length and keeps the first N samples, so a
    batch of segments of many lengths needs only a few FFT plans."""
    import torch

    rng = np.random.default_rng(w.seed)
    N = w.n_samples
    T = make_templates(w, rng)
    onsets, labels, amp = spike_trains(w, rng, N)
    gen = torch.Generator(device=device)
    gen.manual_seed(int(w.seed))
    nfft = 1 << max(0, (N - 1).bit_length()) if pow2_fft else N
    gain = torch.from_numpy(_power_gain(float(w.fs), nfft)).to(device)

    def shaped(k):
        x = torch.randn((k, nfft), generator=gen, device=device, dtype=torch.float32)
        X = torch.fft.rfft(x, dim=1)
        X *= gain
        y = torch.fft.irfft(X, n=nfft, dim=1)[:, :N]
        y /= y.std(dim=1, keepdim=True)
        return y

    V = torch.empty((w.n_chans, N), device=device, dtype=torch.float32)
    shared = shaped(1)[0] if w.noise_corr > 0 else None
    for c0 in range(0, w.n_chans, 8):
        c1 = min(w.n_chans, c0 + 8)
        y = shaped(c1 - c0)
        if shared is not None:
            y = np.sqrt(1 - w.noise_corr) * y + np.sqrt(w.noise_corr) * shared[None, :]
        V[c0:c1] = y
        del y
    Tf = torch.from_numpy(T.astype(np.float32)).to(device)
    U, C, L = T.shape
    flat = V.view(-1)
    ar = torch.arange(L, device=device)
    on_t = torch.from_numpy(onsets).to(device)
    lab_t = torch.from_numpy(labels).to(device)
    amp_t = torch.from_numpy(amp.astype(np.float32)).to(device)
    nz = torch.from_numpy(np.any(T != 0, axis=2)).to(device)  # [U][C] channels each template touches
    # every (spike, touched channel) row at once, in slices of spikes to bound the
    # index tensors; the deterministic accumulation makes the recording the same
    # bits on every run although spikes overlap
    step = max(1, (1 << 24) // max(1, C * L))
    was_det = torch.are_deterministic_algorithms_enabled()
    torch.use_deterministic_algorithms(True)
    try:
        for k0 in range(0, len(onsets), step):
            lab_k = lab_t[k0:k0 + step]
            ks, cs = nz[lab_k].nonzero(as_tuple=True)
            if ks.numel() == 0:
                continue
            idx = ((cs * N + on_t[k0:k0 + step][ks])[:, None] + ar[None, :]).reshape(-1)
            val = (amp_t[k0:k0 + step][ks][:, None] * Tf[lab_k[ks], cs]).reshape(-1)
            flat.index_put_((idx,), val, accumulate=True)
    finally:
        torch.use_deterministic_algorithms(was_det)
    if V.is_cuda:
        torch.cuda.synchronize()
    return package(w, V, T, onsets, labels)
