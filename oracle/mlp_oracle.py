"""CPU float64 restatement of the pruner MLP -- TEST INFRASTRUCTURE.

Forward, loss and backward of the block-pair pruner network
(40 -> 128 -> 128 -> 128 -> 1, SiLU, weighted logistic loss), in plain
numpy float64, as the checker of the GPU kernels' gradients.  The reference
has no pruner; this oracle is pinned by central finite differences of its own
loss (tests/test_prune_oracle.py).

Parameters: dict with W1 [128, 48], b1 [128], W2 [128, 128], b2, W3, b3,
w4 [128], b4 [1], emb [n_emb, 8].  Inputs: the 48-column feature rows whose
embedding columns (3..19) are filled from ``emb`` by the pair's rows.
"""

import numpy

EMB0, EMB1 = 3, 11


def _sig(z):
    return 1.0 / (1.0 + numpy.exp(-z))


def features(base, emb, rows1, rows2):
    """Feature rows: ``base`` [n, 48] with the embedding columns replaced."""
    X = numpy.array(base, dtype=numpy.float64, copy=True)
    X[:, EMB0:EMB0 + 8] = emb[rows1]
    X[:, EMB1:EMB1 + 8] = emb[rows2]
    return X


def forward(params, X):
    """Returns (logits [n], cache)."""
    Z1 = X @ params["W1"].T + params["b1"]
    H1 = Z1 * _sig(Z1)
    Z2 = H1 @ params["W2"].T + params["b2"]
    H2 = Z2 * _sig(Z2)
    Z3 = H2 @ params["W3"].T + params["b3"]
    H3 = Z3 * _sig(Z3)
    logit = H3 @ params["w4"] + params["b4"][0]
    return logit, (X, Z1, H1, Z2, H2, Z3, H3)


def loss_and_grads(params, base, rows1, rows2, labels, pos_weight=1.0):
    """Weighted logistic loss (mean over rows, positives weighted by
    ``pos_weight``) and the gradients of every parameter."""
    X = features(base, params["emb"], rows1, rows2)
    logit, (X, Z1, H1, Z2, H2, Z3, H3) = forward(params, X)
    y = numpy.asarray(labels, dtype=numpy.float64)
    n = X.shape[0]
    sp_neg = numpy.logaddexp(0.0, -logit)
    sp_pos = numpy.logaddexp(0.0, logit)
    loss = float((pos_weight * y * sp_neg + (1 - y) * sp_pos).sum())
    p = _sig(logit)
    dl = (pos_weight * y * (p - 1) + (1 - y) * p) / n
    g = {}
    g["w4"] = H3.T @ dl
    g["b4"] = numpy.array([dl.sum()])
    dsilu = lambda z: _sig(z) * (1 + z * (1 - _sig(z)))  # noqa: E731
    dZ3 = numpy.outer(dl, params["w4"]) * dsilu(Z3)
    g["W3"] = dZ3.T @ H2
    g["b3"] = dZ3.sum(0)
    dZ2 = (dZ3 @ params["W3"]) * dsilu(Z2)
    g["W2"] = dZ2.T @ H1
    g["b2"] = dZ2.sum(0)
    dZ1 = (dZ2 @ params["W2"]) * dsilu(Z1)
    g["W1"] = dZ1.T @ X
    g["b1"] = dZ1.sum(0)
    dX = dZ1 @ params["W1"]
    demb = numpy.zeros_like(params["emb"])
    numpy.add.at(demb, rows1, dX[:, EMB0:EMB0 + 8])
    numpy.add.at(demb, rows2, dX[:, EMB1:EMB1 + 8])
    g["emb"] = demb
    return loss, g, logit
