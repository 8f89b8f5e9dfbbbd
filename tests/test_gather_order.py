"""Gather order of the Delta pass (csrc/host/order.cpp), on the CPU: every
short row exactly once and no long row, whatever the partner lists say;
rows that share columns end up next to one another."""
import numpy as np
import pytest
import torch

from cbm import structure

LONG = structure.LONG_ROW


def _ptr(lengths):
    ptr = np.zeros(len(lengths) + 1, np.int32)
    ptr[1:] = np.cumsum(lengths)
    return torch.from_numpy(ptr)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_every_short_row_once_whatever_the_partners(seed):
    rng = np.random.default_rng(seed)
    n = 400
    lengths = rng.integers(0, 60, n)
    lengths[[3, 77, 250]] = LONG + 1 + rng.integers(0, 40, 3)
    lengths[10] = LONG  # exactly the limit: still short
    ptr = _ptr(lengths)
    # Arbitrary partner lists: out of range, self, long-row and repeated partners.
    partner = torch.from_numpy(rng.integers(-3, n + 3, size=(n, 2)).astype(np.int32))
    partner[5, 0] = 5
    partner[6, 1] = 3
    partner[7] = torch.tensor([8, 8], dtype=torch.int32)
    overlap = torch.from_numpy(rng.integers(0, 40, size=(n, 2)).astype(np.int32))
    order = structure.gather_order(ptr, partner, overlap).numpy()
    short = np.nonzero(lengths <= LONG)[0]
    assert sorted(order.tolist()) == short.tolist()


def test_no_partners_gives_position_order():
    lengths = [3, 0, LONG + 5, 7, 1]
    order = structure.gather_order(_ptr(lengths), torch.full((5, 2), -1, dtype=torch.int32),
                                   torch.zeros((5, 2), dtype=torch.int32)).numpy()
    assert order.tolist() == [0, 1, 3, 4]


def test_walk_follows_the_heaviest_links():
    """A chain 0 - 4 - 2 (heavy links) and 1 - 3 (light): the walk from row 0
    lists 4 and 2 right after it, then starts over at row 1."""
    partner = torch.tensor([[4, -1], [3, -1], [4, -1], [1, -1], [0, 2]], dtype=torch.int32)
    overlap = torch.tensor([[9, 0], [2, 0], [8, 0], [2, 0], [9, 8]], dtype=torch.int32)
    order = structure.gather_order(_ptr([5] * 5), partner, overlap).numpy()
    assert order.tolist() == [0, 4, 2, 1, 3]


def test_rows_sharing_columns_sit_together():
    """Eight families of eight near-copies of one random row, shuffled; with
    each row's true best two partners, every family occupies one stretch of
    at most two pieces of the order (the best-two links may leave a family in
    two parts)."""
    rng = np.random.default_rng(7)
    fam, per, n_cols = 8, 8, 5000
    rows, family = [], []
    for f in range(fam):
        base = rng.choice(n_cols, 40, replace=False)
        for _ in range(per):
            c = base.copy()
            c[rng.integers(0, 40, 3)] = rng.choice(n_cols, 3)
            rows.append(set(c.tolist()))
            family.append(f)
    perm = rng.permutation(len(rows))
    rows = [rows[i] for i in perm]
    family = np.array(family)[perm]
    n = len(rows)
    partner = np.full((n, 2), -1, np.int32)
    overlap = np.zeros((n, 2), np.int32)
    for r in range(n):
        best = sorted(((len(rows[r] & rows[q]), -q) for q in range(n) if q != r), reverse=True)[:2]
        for j, (o, mq) in enumerate(best):
            partner[r, j], overlap[r, j] = -mq, o
    order = structure.gather_order(_ptr([len(r) for r in rows]), torch.from_numpy(partner),
                                   torch.from_numpy(overlap)).numpy()
    fam_seq = family[order]
    for f in range(fam):
        idx = np.nonzero(fam_seq == f)[0]
        pieces = 1 + int(np.sum(np.diff(idx) > 1))
        assert pieces <= 2


def test_invalid_arguments_are_refused():
    with pytest.raises(Exception):
        structure.gather_order(torch.tensor([1, 2], dtype=torch.int32), torch.zeros((1, 2), dtype=torch.int32),
                               torch.zeros((1, 2), dtype=torch.int32))
