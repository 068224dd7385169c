"""The plan of the block-parallel decision stage (fdvg_pdecide_plan, a host
function of the library; no device needed).  The stage cannot wait on a
workgroup that has not started only if:
- every replay unit (one MB row) comes after the unit of the row above it in
  the same sequence (the replays take their units by ticket in this order);
- every MB's left, top and top-left neighbours come earlier in the cost and
  coding order (the replay decides in that order, an intra MB waits for their
  reconstruction);
and it codes every MB exactly once.  The cost and coding kernels take their
MBs from eight per-XCD lists (fdvg_xcd_order): the batch's MB rows (sequence
after sequence) are dealt to the lists in turn, each list keeps the order
above, and together they hold every MB once.  Checked here on the measured workloads' shapes and on
ragged ones, with planted mistakes that the checks must catch."""
import ctypes as C

import numpy as np
import pytest

from fdv_amd import gpu as G

ROWS = 1          # rows per replay unit


class I2(C.Structure):
    _fields_ = [("x", C.c_int), ("y", C.c_int)]


def plan(sizes):
    lib = G.lib()
    f = lib.fdvg_pdecide_plan
    f.restype = C.c_int
    n = len(sizes)
    w = (C.c_int * n)(*[s[0] for s in sizes])
    h = (C.c_int * n)(*[s[1] for s in sizes])
    mb0, tot = [], 0
    for sw, sh in sizes:
        mb0.append(tot)
        tot += (sw // 16) * (sh // 16)
    units = (I2 * (sum(sh // 16 for _, sh in sizes) + n))()
    corder = (C.c_uint * tot)()
    nunit = f(w, h, (C.c_int * n)(*mb0), n, units, corder)
    return nunit, [(u.x, u.y) for u in units[:max(nunit, 0)]], np.array(corder[:], dtype=np.int64), mb0


def xcd_lists(sizes, corder, mb0):
    lib = G.lib()
    f = lib.fdvg_xcd_order
    f.restype = C.c_int
    n, tot = len(sizes), len(corder)
    xorder = (C.c_uint * tot)()
    xoff = (C.c_int * 9)()
    xmax = f((C.c_int * n)(*[s[0] for s in sizes]), (C.c_int * n)(*[s[1] for s in sizes]), (C.c_int * n)(*mb0), n,
             (C.c_uint * tot)(*corder.tolist()), tot, xorder, xoff)
    xo = np.array(xorder[:], dtype=np.int64)
    return xmax, [xo[xoff[k]:xoff[k + 1]] for k in range(8)]


def check_xcd(sizes, corder, mb0, xmax, lists):
    tot = len(corder)
    assert sorted(np.concatenate(lists).tolist()) == list(range(tot))
    assert xmax == max(len(x) for x in lists)
    at = np.empty(tot, dtype=np.int64)
    at[corder] = np.arange(tot)
    rowbase = np.cumsum([0] + [sh // 16 for _, sh in sizes])
    for k, lst in enumerate(lists):
        assert (np.diff(at[lst]) > 0).all(), "a list out of the replays' order"
        for mb in lst.tolist():
            j = max(j for j in range(len(sizes)) if mb >= mb0[j])
            mw = sizes[j][0] // 16
            assert (rowbase[j] + (mb - mb0[j]) // mw) % 8 == k, "an MB in another row's list"


def check(sizes, nunit, units, corder, mb0):
    # units: each sequence's row groups, each after the group above it
    want = sorted((k, g) for k, (sw, sh) in enumerate(sizes) for g in range((sh // 16 + ROWS - 1) // ROWS))
    assert sorted(units) == want
    pos = {u: i for i, u in enumerate(units)}
    for k, g in units:
        if g > 0:
            assert pos[(k, g - 1)] < pos[(k, g)]
    # coding order: a permutation in which every neighbour comes first
    tot = sum((sw // 16) * (sh // 16) for sw, sh in sizes)
    assert sorted(corder.tolist()) == list(range(tot))
    at = np.empty(tot, dtype=np.int64)
    at[corder] = np.arange(tot)
    for k, (sw, sh) in enumerate(sizes):
        mw, mh = sw // 16, sh // 16
        idx = mb0[k] + np.arange(mw * mh).reshape(mh, mw)
        a = at[idx]
        assert (a[:, 1:] > a[:, :-1]).all(), "left neighbour later"
        assert (a[1:, :] > a[:-1, :]).all(), "top neighbour later"
        assert (a[1:, 1:] > a[:-1, :-1]).all(), "top-left neighbour later"


def bands(w, h, rows):
    out, y = [], 0
    while y < h:
        bh = min(rows * 16, h - y)
        out.append((w, bh))
        y += bh
    return out


CASES = {
    "720p eight bands": bands(1280, 720, 6),
    "1088p twelve bands": bands(1920, 1088, 6),
    "16 streams": [(640, 368)] * 8 + [(1280, 720)] * 4 + [(1920, 1088)] * 4,
    "4K bands": bands(3840, 2160, 6),
    "ragged": [(16, 16), (48, 80), (4096, 32), (272, 16 * 9), (64, 16 * 13)],
}


@pytest.mark.parametrize("name", list(CASES))
def test_plan_orders_every_dependency_first(name):
    sizes = CASES[name]
    nunit, units, corder, mb0 = plan(sizes)
    assert nunit == len(units) > 0
    check(sizes, nunit, units, corder, mb0)
    check_xcd(sizes, corder, mb0, *xcd_lists(sizes, corder, mb0))


def test_xcd_check_catches_misplaced_and_reordered_mbs():
    sizes = CASES["720p eight bands"]
    _, _, corder, mb0 = plan(sizes)
    xmax, lists = xcd_lists(sizes, corder, mb0)
    swapped = [x.copy() for x in lists]
    swapped[3][[0, 1]] = swapped[3][[1, 0]]
    with pytest.raises(AssertionError):
        check_xcd(sizes, corder, mb0, xmax, swapped)
    moved = [x.copy() for x in lists]
    moved[0], moved[1] = np.append(moved[0], moved[1][-1]), moved[1][:-1]
    with pytest.raises(AssertionError):
        check_xcd(sizes, corder, mb0, max(len(x) for x in moved), moved)


def test_plan_check_catches_a_swapped_pair():
    sizes = CASES["720p eight bands"]
    nunit, units, corder, mb0 = plan(sizes)
    bad = corder.copy()
    i = int(np.where(bad == mb0[0] + 1)[0][0])        # MB (1, 0) of the first band...
    j = int(np.where(bad == mb0[0])[0][0])            # ...before its left neighbour
    bad[i], bad[j] = bad[j], bad[i]
    with pytest.raises(AssertionError):
        check(sizes, nunit, units, bad, mb0)


def test_plan_of_wide_rows_and_inconsistent_sizes():
    """Rows of any width get a replay unit each; a sequence whose first MB
    index does not follow the one before is refused."""
    sizes = [(4112, 32), (64, 48)]
    nunit, units, corder, mb0 = plan(sizes)
    assert nunit == 5
    check(sizes, nunit, units, corder, mb0)
    lib = G.lib()
    w = (C.c_int * 2)(64, 64)
    h = (C.c_int * 2)(32, 32)
    units_b = (I2 * 8)()
    corder_b = (C.c_uint * 16)()
    assert lib.fdvg_pdecide_plan(w, h, (C.c_int * 2)(0, 7), 2, units_b, corder_b) == -1
