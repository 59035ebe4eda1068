"""Row search (paces/aux/cupy_search.py:16-141): lexicographic lower_bound on packed rows."""

from ..core import engine


def searchsorted_full_binary(phonebook, findme, allow_escapes=False):
    """Index of each row of findme in the sorted phonebook (cupy_search.py:78-93).

    The result is the lexicographic lower_bound in [0, M]; with allow_escapes=False a row that is
    absent raises ValueError.
    """
    if phonebook.shape[1] != findme.shape[1]:
        raise ValueError("Incompatible array shapes.")
    if phonebook.shape[0] >= 2 ** 31 - 1 or findme.shape[0] >= 2 ** 32 - 2:
        raise IndexError("Array too long for 32-bit indices.")
    idx, _, miss = engine.lower_bound(phonebook, findme, want_hit=False,
                                      count_misses=not allow_escapes)
    if not allow_escapes and findme.shape[0] and int(miss.item()):
        raise ValueError("Target values not found in phonebook.")
    return idx
