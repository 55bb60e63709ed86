"""Entropy S(E) and heat capacity Cv(T) from walker state.

Restates the reference's offline analysis:

- `walker_entropy` follows `plotting/parse-binning.py:150-171`
  (`MC.excess_entropy`): for SAD, ln w on [too_lo, too_hi] as it stands;
  below too_lo, ln w(i_lo) + (E - too_lo)/min_T + ln(H/<H>) where <H> is the
  mean histogram over [too_lo, too_hi]; above too_hi, ln w(i_hi) + ln(H/<H>);
  NaN and infinite entries (unvisited bins) become 0; the result is shifted so
  its maximum is 0.  For WL with gamma = 0 (production), ln of the WL
  histogram is added where it is positive.
- `heat_capacity` follows `plotting/final_heat_capacity.py:81-89`:
  Cv(T) = <(E - U)^2> / T^2 under P(E) proportional to exp(S(E) - E/T).
- `mean_entropy` combines many SAD walkers (the reference analyses one
  walker per file): every walker's S is placed on a common energy grid by
  bin index, its free constant fixed by its mean over a common energy window,
  and the estimates are averaged over the walkers that cover each bin; bins
  outside a walker's [too_lo, too_hi] with an empty histogram carry no
  estimate (the reference's script writes 0 there).  This is the host
  statement of the device reduction sadmc_mean_entropy.

All arithmetic is float64 numpy on data already downloaded from the device.
"""
import numpy as np


def bin_centres(state):
    n = int(state["len"])
    return state["bins_min"] + (np.arange(n) + 0.5) * state["bins_width"]


def walker_entropy(state, bins, min_T=None):
    """(E, S) of one walker from its state and logical bins (parse-binning.py:150-171).

    `bins` is the structured array `Context.bins` returns; `min_T` defaults to
    the walker's own (SAD).
    """
    E = bin_centres(state)
    lnw = np.array(bins["lnw"], dtype=np.float64, copy=True)
    hist = np.asarray(bins["histogram"], dtype=np.float64)
    method = int(state["method"])
    with np.errstate(divide="ignore", invalid="ignore"):
        if method == 0:  # SAD
            too_lo, too_hi = float(state["too_lo"]), float(state["too_hi"])
            mT = float(state["min_T"]) if min_T is None else float(min_T)
            i_lo = int(np.abs(E - too_lo).argmin())
            i_hi = int(np.abs(E - too_hi).argmin())
            mean_hist = hist[i_lo:i_hi + 1].mean()
            lo = E < too_lo
            hi = E > too_hi
            lnw[lo] = lnw[i_lo] + (E[lo] - too_lo) / mT + np.log(hist[lo] / mean_hist)
            lnw[hi] = lnw[i_hi] + np.log(hist[hi] / mean_hist)
            lnw[np.isnan(lnw)] = 0
            lnw[np.isinf(lnw)] = 0
        elif method in (2, 3) and float(state["wl_gamma"]) == 0.0:
            wh = np.asarray(bins["wl_hist"], dtype=np.float64)
            lnw[wh > 0] += np.log(wh[wh > 0])
    return E, lnw - lnw.max()


def heat_capacity(T, E, S):
    """Cv(T) for each temperature in T (final_heat_capacity.py:81-89)."""
    T = np.atleast_1d(np.asarray(T, dtype=np.float64))
    E = np.asarray(E, dtype=np.float64)
    S = np.asarray(S, dtype=np.float64)
    C = np.zeros_like(T)
    for i, t in enumerate(T):
        a = S - E / t
        P = np.exp(a - a.max())
        P = P / P.sum()
        U = (E * P).sum()
        C[i] = ((E - U) ** 2 * P).sum() / t ** 2
    return C


def walker_entropy_masked(state, bins):
    """(E, S, valid) of one SAD walker as the walker-set reduction uses it:
    the parse-binning.py:150-171 extension, with bins outside [too_lo,
    too_hi] whose histogram is 0 marked invalid (no estimate) instead of 0."""
    E = bin_centres(state)
    lnw = np.array(bins["lnw"], dtype=np.float64, copy=True)
    hist = np.asarray(bins["histogram"], dtype=np.float64)
    too_lo, too_hi, mT = float(state["too_lo"]), float(state["too_hi"]), float(state["min_T"])
    i_lo = int(np.abs(E - too_lo).argmin())
    i_hi = int(np.abs(E - too_hi).argmin())
    mean_hist = hist[i_lo:i_hi + 1].mean()
    lo, hi = E < too_lo, E > too_hi
    valid = ~((lo | hi) & (hist == 0))
    with np.errstate(divide="ignore"):
        lnw[lo] = lnw[i_lo] + (E[lo] - too_lo) / mT + np.log(hist[lo] / mean_hist)
        lnw[hi] = lnw[i_hi] + np.log(hist[hi] / mean_hist)
    return E, lnw, valid


def mean_entropy(walkers, window=None):
    """Average S(E) over SAD walkers on a common grid.

    `walkers` is an iterable of (state, bins) pairs.  Each walker's S is
    defined up to a constant, fixed by making its mean over the energy window
    (default: the intersection of the walkers' [too_lo, too_hi]) zero;
    walkers without an estimate in every bin of the window are left out.  Returns (E, S,
    count): bin centres, mean entropy, and how many walkers gave an estimate
    in each bin.  Bins without an estimate are dropped.
    """
    walkers = [(st, b) for st, b in walkers if int(st["status"]) == 0 and int(st["len"]) > 0]
    if not walkers:
        raise ValueError("no walkers")
    w = float(walkers[0][0]["bins_width"])
    gmin = min(float(st["bins_min"]) for st, _ in walkers)
    if window is None:
        window = (max(float(st["too_lo"]) for st, _ in walkers), min(float(st["too_hi"]) for st, _ in walkers))
    elif not window[0] < window[1]:
        raise ValueError("empty energy window")
    ga = int(np.ceil((window[0] - gmin) / w - 0.5))
    gb = int(np.floor((window[1] - gmin) / w - 0.5)) + 1
    if gb <= ga:
        raise ValueError("no common energy window for the walkers")
    offs = [int(round((float(st["bins_min"]) - gmin) / w)) for st, _ in walkers]
    n = max(o + int(st["len"]) for o, (st, _) in zip(offs, walkers))
    total = np.zeros(n)
    count = np.zeros(n, dtype=np.int64)
    for o, (st, b) in zip(offs, walkers):
        _, S, valid = walker_entropy_masked(st, b)
        g = o + np.arange(len(S))
        inwin = (g >= ga) & (g < gb)
        if inwin.sum() != gb - ga or not valid[inwin].all():
            continue
        S = S - S[inwin].mean()
        total[g[valid]] += S[valid]
        count[g[valid]] += 1
    keep = count > 0
    E = gmin + (np.arange(n) + 0.5) * w
    return E[keep], total[keep] / count[keep], count[keep]
