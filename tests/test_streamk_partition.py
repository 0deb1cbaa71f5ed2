"""Bookkeeping of the three-stage kernel's batch tail (host-side model).

The tail rows (fewer than one round of workgroups) are split evenly over G
workgroups as (row, block) units; each share writes a row it holds whole to
the row's own slot and a row its start or end cuts to share slot 0 or 1, and
the merge reads the parts back in share order.  The model below walks the
same partition the kernels use (ugsdr_amd/csrc/corr_common.hpp sk_begin /
sk_owner, corr1023.hip corr1023_sk / corr1023_merge_row) and checks the
invariants that make it correct for any batch: every unit is computed
exactly once, no slot is written twice, and the merge reads exactly the
slots that hold the row's parts.
"""
import random

import pytest


def sk_begin(units, G, w):
    return units * w // G


def sk_owner(units, G, x):
    return ((x + 1) * G - 1) // units


def tail_model(Ks, rows_per_req, full_rows):
    """requests (K, rows) laid out in order; rows >= full_rows form the tail"""
    reqs, row, units = [], 0, 0
    for K, n in zip(Ks, rows_per_req):
        a0 = max(row, full_rows) - row
        tail_row0 = min(a0, n)
        reqs.append(dict(K=K, row_start=row, n=n, tail_row0=tail_row0, tail_unit_start=units))
        units += (n - tail_row0) * K
        row += n
    return reqs, row, units


def find_req_unit(reqs, u):
    r = 0
    while r + 1 < len(reqs) and reqs[r + 1]["tail_unit_start"] <= u:
        r += 1
    return r


def find_req_row(reqs, row):
    r = 0
    while r + 1 < len(reqs) and reqs[r + 1]["row_start"] <= row:
        r += 1
    return r


def simulate(Ks, rows_per_req, G_full):
    maxK = max(Ks)
    rows = sum(rows_per_req)
    full_rows = rows // G_full * G_full if maxK > 1 else rows
    reqs, rows, units = tail_model(Ks, rows_per_req, full_rows)
    tail_rows = rows - full_rows
    if tail_rows == 0:
        return 0
    G = min(G_full, units)
    written = {}                      # slot -> (global row, k0, k1)
    covered = {}                      # (global row, k) -> count
    for w in range(G):
        u_begin, u_end = sk_begin(units, G, w), sk_begin(units, G, w + 1)
        u = u_begin
        while u < u_end:
            q = reqs[find_req_unit(reqs, u)]
            lu = u - q["tail_unit_start"]
            local = q["tail_row0"] + lu // q["K"]
            k0 = lu % q["K"]
            k1 = min(q["K"], k0 + (u_end - u))
            grow = q["row_start"] + local
            assert full_rows <= grow < rows
            for k in range(k0, k1):
                covered[(grow, k)] = covered.get((grow, k), 0) + 1
            whole = k0 == 0 and k1 == q["K"]
            slot = ("row", grow - full_rows) if whole else ("share", w, 0 if u == u_begin else 1)
            assert slot not in written, (slot, written[slot], (grow, k0, k1))
            written[slot] = (grow, k0, k1)
            u += k1 - k0
    # every tail unit exactly once
    for grow in range(full_rows, rows):
        q = reqs[find_req_row(reqs, grow)]
        for k in range(q["K"]):
            assert covered.get((grow, k)) == 1, (grow, k)
    # the merge reads exactly the parts of each tail row, covering all its blocks
    read = set()
    for tail in range(tail_rows):
        grow = full_rows + tail
        q = reqs[find_req_row(reqs, grow)]
        local = grow - q["row_start"]
        ur0 = q["tail_unit_start"] + (local - q["tail_row0"]) * q["K"]
        wa, wb = sk_owner(units, G, ur0), sk_owner(units, G, ur0 + q["K"] - 1)
        blocks = []
        for w in range(wa, wb + 1):
            slot = ("row", tail) if wa == wb else ("share", w, 0 if sk_begin(units, G, w) >= ur0 else 1)
            assert slot in written, (tail, slot)
            g2, k0, k1 = written[slot]
            assert g2 == grow
            blocks += list(range(k0, k1))
            read.add(slot)
        assert sorted(blocks) == list(range(q["K"])), (grow, blocks)
    assert read == set(written)
    return tail_rows


def test_single_request_shapes():
    # L1CA_WIDE: 1312 rows x 10 blocks on 256 workgroups -> 32 tail rows
    assert simulate([10], [1312], 256) == 32
    assert simulate([3], [27], 256) == 27          # every row in the tail
    assert simulate([4], [280], 256) == 24         # one round and a tail
    assert simulate([1], [672], 768) == 0          # one block per row: no tail


def test_mixed_batches_random():
    rng = random.Random(1234)
    for _ in range(3000):
        n_req = rng.randint(1, 16)
        Ks = [rng.choice([1, 1, 2, 3, 4, 5, 10, 17]) for _ in range(n_req)]
        rows = [rng.randint(1, 400) for _ in range(n_req)]
        G = rng.choice([7, 64, 256, 768])
        simulate(Ks, rows, G)


@pytest.mark.parametrize("G", [1, 2, 3, 256])
def test_tiny_grids(G):
    simulate([5, 1, 7], [3, 9, 2], G)
