"""The device BLAS build (kpt_blas.hip k_bb_ranks / k_bb_scatter) replaces
the builder's in-place two-pointer partition (as_builder.cuh:201-211) by a
closed-form permutation computed from one scan of the go-right flags.  The
permutation below is the one the kernels compute, written slot by slot from
the same prefix counts; it must equal the sequential loop on every input."""
import itertools

import numpy as np


def sequential(goes_left):
    """The builder's loop: i from the front, j from the back, a right-goer is
    swapped with slot j and j steps down."""
    a = list(range(len(goes_left)))
    i, j = 0, len(goes_left) - 1
    while i <= j:
        if goes_left[a[i]]:
            i += 1
        else:
            a[i], a[j] = a[j], a[i]
            j -= 1
    return a, i


def closed_form(goes_left):
    """Destination of every slot from prefix counts, as the kernels do it."""
    n = len(goes_left)
    flag = np.array([0 if x else 1 for x in goes_left], np.int64)     # 1 = goes right
    scan = np.concatenate([[0], np.cumsum(flag)])                       # exclusive, n+1 entries
    L = n - int(scan[n])
    A = L + 1 if L < n and flag[L] else L
    fr = np.full(n, -1, np.int64)
    bl = np.full(n, -1, np.int64)
    for r in range(n):                                                  # k_bb_ranks
        rb = int(scan[r])
        if r < A:
            if flag[r]:
                fr[rb] = r
        elif not flag[r]:
            bl[L - (r - rb) - 1] = r
    out = [None] * n
    for r in range(n):                                                  # k_bb_scatter
        rb, f = int(scan[r]), flag[r]
        if r < A:
            dest = r if not f else (n - 1 if rb == 0 else int(bl[rb - 1]) - 1)
        else:
            dest = r - 1 if f else int(fr[L - (r - rb) - 1])
        assert 0 <= dest < n and out[dest] is None
        out[dest] = r
    return out, L


def test_every_pattern_up_to_12_slots():
    for n in range(0, 13):
        for bits in itertools.product([False, True], repeat=n):
            assert closed_form(list(bits)) == sequential(list(bits)), bits


def test_random_patterns():
    rng = np.random.default_rng(11)
    for _ in range(3000):
        n = int(rng.integers(1, 400))
        x = list(rng.random(n) < rng.random())
        assert closed_form(x) == sequential(x)


def test_all_right_is_a_rotation():
    """No lefts: the split aborts after the loop has rotated the slots by one
    (the first slot ends last), which the device build must reproduce."""
    out, L = sequential([False] * 5)
    assert L == 0 and out == [1, 2, 3, 4, 0]
    assert closed_form([False] * 5) == (out, L)
