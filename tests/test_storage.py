"""The on-disk structure format (cbm/storage.py): a saved structure reads
back identical, the file holds no pickles, and damaged or foreign files are
refused (CPU only)."""
import io
import json
import zipfile

import numpy as np
import pytest
import torch

from cbm import storage, structure


@pytest.fixture(scope="module")
def built():
    ei = structure.copying_model(3000, 5, 0.85, 0.5, 0, 77)
    b = structure.build(ei[0], ei[1], torch.ones(ei.shape[1]), 3000, 2)
    pl = structure.plan(3000, b.csr_ptr, b.csr_col, b.csr_val, b.tree_src, b.tree_dst)
    return b, pl


def _save(path, b, pl, **over):
    kw = dict(n=3000, alpha=2, delta_crow=b.csr_ptr, delta_col=b.csr_col, delta_sign=torch.sign(b.csr_val),
              mca_offset=b.branches, mca_src_idx=b.tree_src, mca_dst_idx=b.tree_dst, plan=pl,
              build_stats=b.stats, saved_by="cbm4mm")
    kw.update(over)
    storage.save(path, **kw)


def test_round_trip(built, tmp_path):
    b, pl = built
    path = tmp_path / "m.cbm.npz"
    _save(path, b, pl)
    st = storage.load(path)
    assert st.n == 3000 and st.alpha == 2 and st.saved_by == "cbm4mm"
    assert st.build_stats == b.stats
    assert torch.equal(st.delta_crow, b.csr_ptr) and torch.equal(st.delta_col, b.csr_col)
    assert torch.equal(st.delta_values(), b.csr_val)
    assert torch.equal(st.mca_offset, b.branches)
    assert torch.equal(st.mca_src_idx, b.tree_src) and torch.equal(st.mca_dst_idx, b.tree_dst)
    assert st.plan.depth == pl.depth and st.plan.stats == pl.stats
    for k in structure.PLAN_NAMES:
        assert torch.equal(st.plan.arrays[k], pl.arrays[k]), k
    # nothing in the file needs unpickling
    with np.load(path, allow_pickle=False) as z:
        for name in z.files:
            assert z[name].dtype != object


def _rewrite(src, dst, edit):
    """Copies a structure file, passing each member's array through edit(name, array)."""
    with np.load(src, allow_pickle=False) as z:
        arrays = {k: z[k] for k in z.files}
    arrays = {k: edit(k, v) for k, v in arrays.items()}
    arrays = {k: v for k, v in arrays.items() if v is not None}
    with open(dst, "wb") as fh:
        np.savez(fh, **arrays)


def test_damaged_files_are_refused(built, tmp_path):
    b, pl = built
    good = tmp_path / "good.npz"
    _save(good, b, pl)

    def flip(name, a):
        if name == "delta_col":
            a = a.copy()
            a[17] ^= 1
        return a
    _rewrite(good, tmp_path / "flip.npz", flip)
    with pytest.raises(storage.StructureFileError, match="checksum"):
        storage.load(tmp_path / "flip.npz")

    _rewrite(good, tmp_path / "missing.npz", lambda k, a: None if k == "plan/bwd_row" else a)
    with pytest.raises(storage.StructureFileError, match="missing"):
        storage.load(tmp_path / "missing.npz")

    def version(name, a):
        if name == "meta":
            m = json.loads(bytes(a).decode())
            m["version"] = 99
            return np.frombuffer(json.dumps(m).encode(), dtype=np.uint8)
        return a
    _rewrite(good, tmp_path / "version.npz", version)
    with pytest.raises(storage.StructureFileError, match="version"):
        storage.load(tmp_path / "version.npz")

    _rewrite(good, tmp_path / "dtype.npz", lambda k, a: a.astype(np.int64) if k == "mca_src_idx" else a)
    with pytest.raises(storage.StructureFileError, match="int32"):
        storage.load(tmp_path / "dtype.npz")

    (tmp_path / "junk.npz").write_bytes(b"not a zip")
    with pytest.raises(storage.StructureFileError, match="not a structure file"):
        storage.load(tmp_path / "junk.npz")

    buf = io.BytesIO()
    with zipfile.ZipFile(buf, "w") as zf:
        zf.writestr("x.npy", b"")
    (tmp_path / "nometa.npz").write_bytes(buf.getvalue())
    with pytest.raises(storage.StructureFileError):
        storage.load(tmp_path / "nometa.npz")


def test_inconsistent_sizes_are_refused(built, tmp_path):
    """A file whose arrays checksum but do not fit together (n disagrees)."""
    b, pl = built
    path = tmp_path / "n.npz"
    _save(path, b, pl)

    def bad_n(name, a):
        if name == "meta":
            m = json.loads(bytes(a).decode())
            m["n"] = 2999
            return np.frombuffer(json.dumps(m).encode(), dtype=np.uint8)
        return a
    _rewrite(path, tmp_path / "n2.npz", bad_n)
    with pytest.raises(storage.StructureFileError, match="sizes"):
        storage.load(tmp_path / "n2.npz")


def test_save_checks_its_input(built, tmp_path):
    b, pl = built
    with pytest.raises(ValueError, match="n x n"):
        _save(tmp_path / "x.npz", b, pl, n=2999)
    with pytest.raises(ValueError, match="delta_sign"):
        _save(tmp_path / "y.npz", b, pl, delta_sign=torch.zeros_like(b.csr_val))


@pytest.mark.parametrize("version", [1, 2, 3, 4])
def test_older_files_load_without_the_newer_layout_parts(built, tmp_path, version):
    """A version 1 file (no 4-range piece set, no tree tails, no units), a
    version 2 file (no tails, no units), a version 3 file (no units) or a
    version 4 file (no gather order) loads with those parts empty -- no
    pieces, tails at the depth, i.e. none, no units, no order -- and the rest
    of the layout unchanged."""
    b, pl = built
    path = tmp_path / "v5.npz"
    _save(path, b, pl, gather_order=_short_rows(pl))
    gone = ((set(storage._NEW_IN_4) if version <= 3 else set()) | (set(storage._NEW_IN_3) if version <= 2 else set())
            | (set(storage._NEW_IN_2) if version == 1 else set()))
    gone_stats = ((set(storage._NEW_STATS_IN_4) if version <= 3 else set())
                  | (set(storage._NEW_STATS_IN_3) if version <= 2 else set())
                  | (set(storage._NEW_STATS_IN_2) if version == 1 else set()))

    def older(name, a):
        if (name.startswith("plan/") and name[5:] in gone) or name == "gather_order":
            return None
        if name == "meta":
            m = json.loads(bytes(a).decode())
            m["version"] = version
            for k in gone_stats:
                m["plan_stats"].pop(k)
            for k in gone:
                m["crc32"].pop("plan/" + k)
            m["crc32"].pop("gather_order")
            return np.frombuffer(json.dumps(m).encode(), dtype=np.uint8)
        return a
    _rewrite(path, tmp_path / "old.npz", older)
    st = storage.load(tmp_path / "old.npz")
    assert st.gather_order is None
    for k in gone_stats:
        assert st.plan.stats[k] == (pl.depth if k in storage._NEW_STATS_IN_3 else 0), k
    for k in structure.PLAN_NAMES:
        if k in gone:
            assert st.plan.arrays[k].numel() == 0, k
        else:
            assert torch.equal(st.plan.arrays[k], pl.arrays[k]), k


def _short_rows(pl):
    return torch.nonzero(pl.arrays["fwd_ptr"].diff() <= structure.LONG_ROW).view(-1).to(torch.int32).flip(0)


def test_gather_order_round_trip(built, tmp_path):
    b, pl = built
    order = _short_rows(pl)
    _save(tmp_path / "o.npz", b, pl, gather_order=order)
    assert torch.equal(storage.load(tmp_path / "o.npz").gather_order, order)
    _save(tmp_path / "none.npz", b, pl)
    assert storage.load(tmp_path / "none.npz").gather_order is None


def test_a_gather_order_of_other_rows_is_refused(built, tmp_path):
    b, pl = built
    order = _short_rows(pl).clone()
    order[0] = order[1]
    _save(tmp_path / "bad.npz", b, pl, gather_order=order)
    with pytest.raises(storage.StructureFileError, match="gather order"):
        storage.load(tmp_path / "bad.npz")
