import numpy as np


def canon(*cols):
    """Canonical order of a multiset of rows: lexicographic sort by all columns."""
    order = np.lexsort(tuple(reversed([np.asarray(c) for c in cols])))
    return tuple(np.asarray(c)[order] for c in cols)


def same_multiset(a_cols, b_cols) -> bool:
    if len(a_cols[0]) != len(b_cols[0]):
        return False
    return all(np.array_equal(x, y) for x, y in zip(canon(*a_cols), canon(*b_cols)))
