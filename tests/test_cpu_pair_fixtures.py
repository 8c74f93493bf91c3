"""Pair lists against the reference's own output (tests/golden/pair_lists.npz, written by
tests/golden/make_pair_list_fixtures.py from reference pair_utility.py:6-133 on the same seeded
molecules): the product's builder (native pre-screen + exact numpy bound) and the oracle's scalar
restatement must both reproduce the reference's class order, pair order (ties included), membership and
bound bits.  The plain case is the full 80-atom workload molecule (308 205 pairs)."""

import pathlib

import numpy as np
import pytest

from electron_integral_playground_amd import generators, pair_utility
from oracle import pair_list as oracle_pairs

FIXTURE = pathlib.Path(__file__).resolve().parent / "golden" / "pair_lists.npz"
CASES = {"small": ("small", None, None), "plain": ("plain", None, None),
         "mixed30": ("mixed", 30, 20.0 * (30 / 200) ** (1 / 3))}
FIELDS = ("i_shell", "j_shell", "i_primitive", "j_primitive", "upper_bound")


def _reference(key):
    z = np.load(FIXTURE)
    classes = [tuple(int(v) for v in c) for c in z[f"{key}/classes"]]
    return classes, {c: {f: z[f"{key}/{c[0]}{c[1]}/{f}"] for f in FIELDS} for c in classes}


def _molecule(key):
    workload, n_atom, box = CASES[key]
    return generators.make_molecule(workload, n_atom=n_atom, box=box)


def _assert_same(got_classes, got, ref_classes, ref, who):
    assert got_classes == ref_classes, f"{who}: class order {got_classes} vs reference {ref_classes}"
    for c in ref_classes:
        for f in FIELDS:
            g, r = np.asarray(got[c][f]), ref[c][f]
            assert g.shape == r.shape, f"{who} {c} {f}: {g.shape} vs {r.shape}"
            if f == "upper_bound":
                assert np.array_equal(g.view(np.int64), r.view(np.int64)), f"{who} {c}: bound bits differ"
            else:
                assert np.array_equal(g.astype(np.int64), r.astype(np.int64)), f"{who} {c} {f} differs"


@pytest.mark.parametrize("key", list(CASES))
def test_product_pair_arrays_equal_reference(key):
    ref_classes, ref = _reference(key)
    pa = pair_utility.form_pair_arrays(_molecule(key), 1e-14)
    got = {c: {f: getattr(a, f) for f in FIELDS} for c, a in pa.items()}
    _assert_same(list(pa), got, ref_classes, ref, "form_pair_arrays")


@pytest.mark.parametrize("key", ["small", "mixed30", "plain"])
def test_oracle_restatement_equals_reference(key):
    ref_classes, ref = _reference(key)
    pl = oracle_pairs.form_primitive_pair_list(_molecule(key), 1e-14)
    got = {c: {f: np.array([getattr(p, f) for p in v]) for f in FIELDS} for c, v in pl.items()}
    _assert_same(list(pl), got, ref_classes, ref, "oracle pair_list")


def test_fixture_is_not_trivial():
    """The fixture holds ties and near-threshold pairs, so order and membership are really pinned."""
    classes, ref = _reference("plain")
    ub = np.concatenate([ref[c]["upper_bound"] for c in classes])
    assert len(ub) == 308205
    ties = sum(int((np.diff(ref[c]["upper_bound"]) == 0).sum()) for c in classes)
    assert ties > 0
    assert ub.min() >= 1e-14 and ub.min() < 1e-13
