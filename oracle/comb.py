"""Comb population control (test infrastructure only).

Follows afqmc/utilities/populationControl/source/populationControlConfiguration.cpp
and include/populationControl.h.  Pinned by the popCheck literals of
populationControlTest.cpp:12-27, by the reference's b_exact table
(populationControlTest.cpp:29-45: weights 1..100 over 4 chunks, reproduced slot
for slot from uniforms placed inside each tooth's bin) and by invariants (every
slot count is 1 afterwards; a slot with count >= 1 keeps itself); the checks are
in tests/test_oracle_models.py.
"""
import numpy as np


def popCheck(weight):
    """(min, max, mean) with sequential sum (populationControlConfiguration.cpp:9-22)."""
    L = len(weight)
    s = 0.0
    for w in weight:
        s += w
    maxW, minW = 0.0, 1e20
    for w in weight:
        if maxW < w:
            maxW = w
        if minW > w:
            minW = w
    return minW, maxW, s / (L * 1.0)


def popConfiguration(size, weight, u):
    """Stratified comb + reassignment (populationControlConfiguration.cpp:24-101).

    ``u`` holds the L uniforms the comb consumes, one per slot, in slot
    order.  The bin search of the reference has no bounds guard; when the
    rounded cumulative sum ends below (L-1+u)/L the index is clamped to L-1
    here (the reference would read past the array)."""
    weight = [float(w) for w in weight]
    L = len(weight)
    L_chunk = L // size
    if L_chunk * size != L:
        raise ValueError("Size of weight can not be divided by Size")
    s = 0.0
    for w in weight:
        s += w
    # IEEE division as in C++: a zero total (every walker dead) gives NaN probabilities, and every
    # tooth then stays in bin 0 (NaN comparisons are false)
    with np.errstate(invalid="ignore", divide="ignore"):
        prob = [float(x) for x in np.array(weight, dtype=np.float64) / np.float64(s)]
    number = [0] * L
    index_old = 0
    prob_sum_old = prob[0]
    for i in range(L):
        prob_sum_new = (i + float(u[i])) / (L * 1.0)
        while prob_sum_new > prob_sum_old:
            if index_old + 1 >= L:
                break
            index_old += 1
            prob_sum_old += prob[index_old]
        number[index_old] += 1
    table = [0] * L
    for i in range(L):
        if number[i] != 0:
            table[i] = i

    def exchange(lo, hi):
        zero_i = lo
        one_i = lo
        while zero_i < hi and number[zero_i] != 0:
            zero_i += 1
        while one_i < hi and number[one_i] <= 1:
            one_i += 1
        while zero_i < hi and one_i < hi:
            table[zero_i] = one_i
            number[zero_i] += 1
            number[one_i] -= 1
            while zero_i < hi and number[zero_i] != 0:
                zero_i += 1
            while one_i < hi and number[one_i] <= 1:
                one_i += 1

    for r in range(size):
        exchange(L_chunk * r, L_chunk * (r + 1))
    if size > 1:
        exchange(0, L)
    for i in range(L):
        if number[i] != 1:
            raise RuntimeError("After pop configuration, number of each walker is not 1!")
    return np.array(table, dtype=np.int32)


def populationControl(walkers, table):
    """Serial branch: walker[i] = walker[table[i]] in slot order (populationControl.h:110-116)."""
    for i in range(len(table)):
        j = int(table[i])
        if i != j:
            walkers[i] = walkers[j].copy()
    return walkers
