"""Drop-in for the reference's ``triangular`` (caller/native/caller/triangular.py):
unordered diploid likelihoods in VCF order (b = 0..n-1, a = 0..b).

:func:`diploid_likelihood` dispatches as triangular.py:6-11 (two columns with
a prior -> the two-allele mixture, else the general one).  The reference's
``_diploid_likelihood_numpy`` (triangular.py:13-24), its canonical CPU form
used by partial_anchor.py:116-117, computes the same mixture; here it runs the
same device mixture, FP64 throughout.
"""
from __future__ import annotations

from .mixture_cpu import general, two

# artifact.py:12-15
BALANCE = (.2, .3, .4, .5, .6, .7, .8)
BALANCE_WEIGHT = (.025, .1, .225, .3, .225, .1, .025)
CONTAMINATION = (.02, .1, .2, .3)
CONTAMINATION_WEIGHT = (.4, .35, .2, .05)


def diploid_likelihood(lh, artifact_prior=None):
    if lh.shape[1] == 2 and artifact_prior is not None:
        return two(lh, artifact_prior).ravel()[[0, 1, 3]]
    return general(lh, artifact_prior, triangular=True)


def _diploid_likelihood_numpy(lh, artifact_prior=None):
    return general(lh, artifact_prior, triangular=True)


def pair_order(n):
    """The (a, b) of every output index: triangular.py:14."""
    return [(a, b) for b in range(n) for a in range(b + 1)]
