"""Drop-in for the reference's ``triangular`` (caller/native/caller/triangular.py):
unordered diploid likelihoods in VCF order (b = 0..n-1, a = 0..b).

* :func:`diploid_likelihood` dispatches as triangular.py:6-11: two haplotype
  columns with an artifact prior go to the two-allele mixture
  (mixture_two.cpp), anything else to the general one (mixture_general.cpp),
  both on the device in FP64.
* :func:`_diploid_likelihood_numpy` is the name partial_anchor.py:116-117
  imports for its canonical recompute of a near-tied locus.  The reference's
  body (triangular.py:13-24) is a numpy evaluation of the same mixture as
  mixture_general.cpp with the same operation order -- per pair and mixture
  term log(max(1e-300, w a + (1-w) b)) summed over rows in row order, the
  balance terms combined by log-add in index order, the artifact terms
  theta-major, then log1p(-prior) + clean log-added to log(prior) + art -- so
  here it is the general mixture on the device (FP64, triangular order).
  The two forms differ only in the last bits of their log implementations
  (pinned in tests/test_oracle_genotype.py to <= 1e-13 relative).
* :func:`pair_order` lists the (a, b) of every output index.
"""
from __future__ import annotations

from .mixture_cpu import general, two

# artifact.py:12-15
BALANCE = (.2, .3, .4, .5, .6, .7, .8)
BALANCE_WEIGHT = (.025, .1, .225, .3, .225, .1, .025)
CONTAMINATION = (.02, .1, .2, .3)
CONTAMINATION_WEIGHT = (.4, .35, .2, .05)


def diploid_likelihood(lh, artifact_prior=None):
    """triangular.py:6-11."""
    if lh.shape[1] == 2 and artifact_prior is not None:
        return two(lh, artifact_prior).ravel()[[0, 1, 3]]
    return general(lh, artifact_prior, triangular=True)


def _diploid_likelihood_numpy(lh, artifact_prior=None):
    """The canonical form of triangular.py:13-24: the general mixture for any
    number of columns (no two-allele dispatch), triangular order."""
    return general(lh, artifact_prior, triangular=True)


def pair_order(n):
    """The (a, b) of every output index: triangular.py:14."""
    return [(a, b) for b in range(n) for a in range(b + 1)]
