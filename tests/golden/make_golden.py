"""Writes tests/golden/cbm_small.json: small binary matrices with the CBM
structure and products they must give, computed by brute force.

The reference ships no fixtures (it has no tests), so these vectors come from
the definitions alone, independently of both the oracle and the product:
  * candidate edges by the rule of cbm/cbm_extensions.cpp:122-182 on dense
    numpy row overlaps, in the reference creation order;
  * the minimum arborescence by exhaustive enumeration of all in-edge
    choices (only instances whose optimum is unique are kept, so the parent
    array is fixed by the mathematics and not by a tie-break);
  * Delta_v = A_v - A_parent(v) as dense vectors;
  * products A X, A D X, D A D X and transposes in exact integer / FP64
    arithmetic with small-integer X.

    python tests/golden/make_golden.py      (rewrites the JSON)
"""
import itertools
import json
import os

import numpy as np

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cbm_small.json")


def candidates(A, alpha):
    n = A.shape[0]
    deg = A.sum(1)
    S = A @ A.T
    out = []
    for r in range(n):
        out.append((n, r, int(deg[r])))
        for c in range(r + 1, n):
            if S[r, c] == 0:
                continue
            h = int(deg[r] + deg[c] - 2 * S[r, c])
            if h < deg[r] - alpha and h != deg[r]:
                out.append((c, r, h))
            if h < deg[c] - alpha and h != deg[c]:
                out.append((r, c, h))
    return out


def brute_force_msa(n, edges):
    """All optimal parent arrays (by enumeration of one in-edge per node)."""
    ins = [[e for e in edges if e[1] == v] for v in range(n)]
    best, sols = None, []
    for choice in itertools.product(*ins):
        parent = [e[0] for e in choice]
        ok = True
        for v in range(n):  # every node must reach the root n
            seen, u = set(), v
            while u != n:
                if u in seen:
                    ok = False
                    break
                seen.add(u)
                u = parent[u]
            if not ok:
                break
        if not ok:
            continue
        w = sum(e[2] for e in choice)
        if best is None or w < best:
            best, sols = w, [parent]
        elif w == best:
            sols.append(parent)
    return best, sols


def random_instance(rng, n, p):
    """Symmetric 0/1 matrix; half the instances copy rows with noise so that
    non-trivial trees (rows close in Hamming distance) are common."""
    A = (rng.random((n, n)) < p).astype(np.int64)
    if rng.random() < 0.5:
        for i in range(1, n):
            if rng.random() < 0.7:
                j = int(rng.integers(0, i))
                A[i] = np.where(rng.random(n) < 0.85, A[j], A[i])
    A = np.triu(A, 1)
    A = A + A.T
    return A


def products(A, X):
    deg = A.sum(1).astype(np.float64)
    d = np.zeros_like(deg)
    d[deg > 0] = deg[deg > 0] ** -0.5
    D = np.diag(d)
    Af = A.astype(np.float64)
    return {"ax": Af @ X, "adx": Af @ D @ X, "dadx": D @ Af @ D @ X,
            "ax_t": Af.T @ X, "adx_t": (Af @ D).T @ X, "dadx_t": (D @ Af @ D).T @ X}


def main():
    rng = np.random.default_rng(20260601)
    cases = []
    tries = 0
    trivial = 0
    while len(cases) < 24 and tries < 20000:
        tries += 1
        n = int(rng.integers(3, 8))
        A = random_instance(rng, n, float(rng.uniform(0.3, 0.8)))
        if A.sum() == 0:
            continue
        alpha = int(rng.integers(0, 2))
        edges = candidates(A, alpha)
        weight, sols = brute_force_msa(n, edges)
        if len(sols) != 1:
            continue
        parent = [(-1 if u == n else int(u)) for u in sols[0]]
        if all(p < 0 for p in parent):
            if trivial >= 4:
                continue
            trivial += 1
        delta = A.copy()
        for v in range(n):
            if parent[v] >= 0:
                delta[v] = A[v] - A[parent[v]]
        X = rng.integers(-4, 5, size=(n, 3)).astype(np.float64)
        r, c = np.nonzero(A)
        cases.append({
            "n": n, "alpha": alpha,
            "edge_index": [r.tolist(), c.tolist()],
            "candidates": [list(e) for e in edges],
            "weight": int(weight),
            "parent": parent,
            "delta_dense": delta.tolist(),
            "x": X.tolist(),
            "products": {k: v.tolist() for k, v in products(A, X).items()},
        })
    with open(OUT, "w") as fh:
        json.dump({"generator": "tests/golden/make_golden.py", "cases": cases}, fh, indent=0)
    print(f"wrote {len(cases)} cases to {OUT} ({tries} random instances tried)")


if __name__ == "__main__":
    main()
