"""Chi-square helpers shared by the sampler tests."""
import numpy as np
from scipy.stats import chisquare


def chi2_pvalue(samples, probs, min_expected=5.0):
    """p-value of samples against probs; outcomes with < min_expected counts are pooled."""
    probs = np.asarray(probs, dtype=np.float64)
    probs = probs / probs.sum()
    k = len(samples)
    counts = np.bincount(np.asarray(samples, dtype=np.int64), minlength=probs.size)
    keep = probs * k >= min_expected
    f_obs = counts[keep].astype(np.float64)
    f_exp = probs[keep] * k
    rest_e = probs[~keep].sum() * k
    rest_o = counts[~keep].sum()
    if rest_e > 0:
        f_obs = np.append(f_obs, rest_o)
        f_exp = np.append(f_exp, rest_e)
    else:
        assert rest_o == 0, "samples fell on zero-probability outcomes"
    f_exp *= f_obs.sum() / f_exp.sum()
    return chisquare(f_obs, f_exp).pvalue


def marginal_pvalues(samples, probs, n):
    """Per-qubit marginal chi-square p-values."""
    probs = np.asarray(probs, dtype=np.float64)
    idx = np.arange(probs.size)
    s = np.asarray(samples, dtype=np.int64)
    out = []
    for q in range(n):
        p1 = probs[((idx >> q) & 1) == 1].sum() / probs.sum()
        out.append(chi2_pvalue((s >> q) & 1, [1 - p1, p1]))
    return out
