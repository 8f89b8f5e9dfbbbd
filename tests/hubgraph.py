"""A dense copying-model graph with three hub nodes: many Delta rows longer
than one workgroup chunk, in both the forward (rows) and backward (columns)
orders.  Shared by the CPU plan tests and the GPU parity tests."""
import numpy as np
import torch

from cbm import structure

HUB_N = 6000


def hub_graph(n=HUB_N, seed=77):
    ei = structure.copying_model(n, 2, 0.95, 3.0, 1500, seed).long().numpy()
    rng = np.random.default_rng(seed)
    parts = [ei]
    for h in (10, 20, 30):
        nb = np.nonzero(rng.random(n) < 0.6)[0]
        nb = nb[nb != h]
        parts.append(np.stack([np.full(len(nb), h), nb]))
        parts.append(np.stack([nb, np.full(len(nb), h)]))
    allr = np.concatenate(parts, axis=1)
    key = np.unique(allr[0] * n + allr[1])
    return torch.from_numpy(np.stack([key // n, key % n]).astype(np.int32))


STAR_K = 2500


def star_graph(k=STAR_K):
    """A prototype with k near-copies: nodes 0..19 form a core B, node 20 is
    the prototype P with neighbours B, node 21 + i is copy c_i with
    neighbours B plus its own pendant node 21 + k + i.  Each c_i is one entry
    away from P and two from any other copy, so the unique cheapest parent of
    every copy is P: P has k children, a child list several chunks long in
    the adjoint's reverse pass (and the B rows are long Delta rows)."""
    n = 21 + 2 * k
    rows, cols = [], []
    core = np.arange(20)
    for c in [20] + list(range(21, 21 + k)):
        rows.append(np.full(20, c))
        cols.append(core)
    copies = np.arange(21, 21 + k)
    rows.append(copies)
    cols.append(copies + k)
    r = np.concatenate(rows)
    c = np.concatenate(cols)
    r, c = np.concatenate([r, c]), np.concatenate([c, r])
    key = np.unique(r * n + c)
    return torch.from_numpy(np.stack([key // n, key % n]).astype(np.int32)), n
