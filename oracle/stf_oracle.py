"""Restatement of stf.c (source time function by Wiener deconvolution) and
of the trace normalisations of util.c:489-579 in numpy.

TEST INFRASTRUCTURE ONLY: checks dn_stf / dn_stf_shot in tests/.

stf.c:38-250 for one shot: killed / muted traces are zeroed, NORMALIZE 1
divides samples 2..ns of each trace by the trace's largest |sample| over
those samples (0 -> 1), NORMALIZE 2 multiplies the observed samples 2..ns by
max_syn / max_obs; with Npad = 2^(ceil(log2 ns) + 2) and FFTs of the
zero-padded traces:
    S_n = sum_i ntr D_i conj(M_i),  S_d = sum_i ntr |M_i|^2,
    Ebar = sum_i sum_k (S_d after trace i)[k] / (Npad ntr)   (float),
    STF = Re IFFT(FFT(s) S_n / (S_d + EPS_STF ntr Ebar))[:ns].
The reference's FFTW calls use sign +1 forward and -1 inverse; for real
input the real part of the result equals the usual convention's (numpy).
Ebar is summed in float64 here (float in the reference).
Pin: with EPS_STF = 0 and observed traces equal to the modelled ones
convolved with a filter h (the convolution fitting the record), the
estimate equals signal * h exactly (tests/test_oracle_pins.py).
"""
import math

import numpy as np


def normalize_max(x):
    """util.c:489-526 on a copy of [ntr][ns]."""
    x = x.copy()
    for i in range(x.shape[0]):
        seg = x[i, 1:]
        mx = max(float(max(seg.max(initial=0.0), 0.0)), abs(float(min(seg.min(initial=0.0), 0.0))))
        if mx == 0.0:
            mx = 1.0
        x[i, 1:] = (seg / x.dtype.type(mx)).astype(x.dtype)
    return x


def trace_max(x):
    out = []
    for i in range(x.shape[0]):
        seg = x[i, 1:]
        mx = max(float(max(seg.max(initial=0.0), 0.0)), abs(float(min(seg.min(initial=0.0), 0.0))))
        out.append(1.0 if mx == 0.0 else mx)
    return np.array(out, x.dtype)


def stf(obs, syn, signal, mute=None, eps_stf=0.05, normalize=0):
    dt = np.float64 if np.asarray(obs).dtype == np.float64 else np.float32
    o = np.array(obs, dt)
    m = np.array(syn, dt)
    ntr, ns = o.shape
    if mute is not None:
        z = np.asarray(mute).astype(bool)
        o[z] = 0
        m[z] = 0
    if normalize == 1:
        m = normalize_max(m)
        o = normalize_max(o)
    elif normalize == 2:
        ratio = (trace_max(m) / trace_max(o)).astype(dt)
        o[:, 1:] = (o[:, 1:] * ratio[:, None]).astype(dt)
    N = 2 ** (int(math.ceil(math.log(ns) / math.log(2.0))) + 2)
    D = np.fft.fft(o.astype(np.float64), N, axis=1)
    M = np.fft.fft(m.astype(np.float64), N, axis=1)
    S = np.fft.fft(np.asarray(signal, np.float64), N)
    n = float(ntr)
    sn = (n * D * np.conj(M)).sum(axis=0)
    p = n * (M.real ** 2 + M.imag ** 2)
    sd = p.sum(axis=0)
    ebar = np.cumsum(p, axis=0).sum() / (N * ntr)
    if dt == np.float32:
        ebar = np.float32(ebar)
        eps = float(np.float32(np.float32(eps_stf) * np.float32(ntr)) * ebar)
    else:
        eps = eps_stf * ntr * ebar
    y = np.fft.ifft(S * sn / (sd + eps)).real[:ns]
    return y.astype(dt)
