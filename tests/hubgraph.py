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
