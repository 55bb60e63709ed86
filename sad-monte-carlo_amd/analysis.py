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
- `mean_entropy` combines many walkers (the reference analyses one walker
  per file): every walker's S is placed on a common energy grid by bin
  index, its free constant fixed on the bins all walkers cover, and the
  estimates are averaged over the walkers that cover each bin.

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


def mean_entropy(walkers, width=None, e_min=None):
    """Average S(E) over walkers on a common grid.

    `walkers` is an iterable of (state, bins) pairs.  Each walker's S is
    defined up to a constant; the constants are fixed by making every walker
    average to zero over the bins that all walkers cover (or, with no such
    bin, by normalising sum exp(S) = 1 per walker).  Returns (E, S, count):
    bin centres, mean entropy (maximum 0), and how many walkers covered each
    bin.  Bins no walker covers are dropped.
    """
    per = []
    for st, b in walkers:
        _, S = walker_entropy(st, b)
        per.append((float(st["bins_min"]), float(st["bins_width"]), S))
    if not per:
        raise ValueError("no walkers")
    w = per[0][1] if width is None else width
    if any(p[1] != w for p in per):
        raise ValueError("walkers with different bin widths")
    lo = min(p[0] for p in per) if e_min is None else e_min
    offs = [int(round((p[0] - lo) / w)) for p in per]
    n = max(o + len(p[2]) for o, p in zip(offs, per))
    count = np.zeros(n, dtype=np.int64)
    for o, p in zip(offs, per):
        count[o:o + len(p[2])] += 1
    common = count == len(per)
    total = np.zeros(n)
    for o, p in zip(offs, per):
        S = p[2]
        mine = common[o:o + len(S)]
        S = S - (S[mine].mean() if mine.any() else np.log(np.exp(S).sum()))
        total[o:o + len(S)] += S
    keep = count > 0
    E = lo + (np.arange(n) + 0.5) * w
    S = total[keep] / count[keep]
    return E[keep], S - S.max(), count[keep]
