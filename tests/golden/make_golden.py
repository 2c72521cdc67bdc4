"""Writes the committed fixtures of tests/golden/ (run from the repository root).

logreg_L2.npz
    For an alignment of L = 2 loci the pseudolikelihood of column r = 0 is a weighted
    multinomial logistic regression of the state at locus 0 on the one-hot state at
    locus 1 plus an intercept (plmDCA_objective.hpp:86-151 with a single j).  With
    lambda_h = lambda_J the L2 terms equal scikit-learn's penalty 1/(2C) ||W||^2 on a
    design matrix [onehot(x_1), 1] without a separate intercept, C = 1 / (2 lambda).
    The fixture holds the inputs and scikit-learn's optimum (coef_ as J^T and h).
    scikit-learn is a library routine independent of this repository and of SuperDCA.

couplings_tiny.all
    Output of the coupling writer for a 4 x 4 score matrix whose values are exactly
    representable, with translation [2, 5, 7, 11] and base 1: the order and text
    format of plmDCA.hpp:667-693 ("%.8f i j", i > j, r in list order, n before r).
"""
from __future__ import annotations

import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def make_logreg():
    from sklearn.linear_model import LogisticRegression

    rng = np.random.default_rng(20240611)
    B, q = 80, 4
    x0 = rng.integers(0, q, B)  # target locus: all four states present
    x0[:q] = np.arange(q)
    x1 = rng.integers(0, q, B)
    x1 = np.where(rng.random(B) < 0.6, x0, x1)  # coupled to x0
    x1[x1 == 2] = 1  # locus 1 observes states {0, 1, 3}: one state never seen
    states = np.stack([x0, x1], axis=1).astype(np.uint8)
    w = rng.uniform(0.2, 1.0, B)
    lam = 1.7  # absolute lambda_h = lambda_J
    X = np.zeros((B, q + 1))
    X[np.arange(B), x1] = 1.0
    X[:, q] = 1.0
    clf = LogisticRegression(C=1.0 / (2.0 * lam), fit_intercept=False, tol=1e-12, max_iter=100000)
    clf.fit(X, x0, sample_weight=w)
    coef = clf.coef_  # q classes (state s at r) x (q + 1) features
    np.savez(os.path.join(HERE, "logreg_L2.npz"), states=states, w=w, lam=np.array(lam),
             J1=coef[:, :q].T.copy(), h=coef[:, q].copy())


def make_writer():
    # Expected text written with Python's own "%.8f" conversion (correctly rounded, as
    # printf's), so the fixture does not come from the writer under test.
    S = np.array([[0.0, 0.25, 0.5, 0.125],
                  [0.75, 0.0, 1.5, 2.0],
                  [0.0625, 0.375, 0.0, 0.03125],
                  [1.0, 0.5, 0.25, 0.0]], dtype=np.float32)
    tr = [2, 5, 7, 11]
    lines = []
    for r in range(4):
        for n in range(r):
            v = float(np.float32(S[r, n] + S[n, r])) * 0.5
            lines.append("%.8f %d %d\n" % (v, tr[r] + 1, tr[n] + 1))
    with open(os.path.join(HERE, "couplings_tiny.all"), "w") as fh:
        fh.writelines(lines)
    np.save(os.path.join(HERE, "couplings_tiny_S.npy"), S)


if __name__ == "__main__":
    make_logreg()
    make_writer()
