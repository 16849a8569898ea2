"""LDS index bounds of the batch kernel's grid transforms, checked on the CPU.

The batch kernel transforms each cell's grid through two LDS buffers of
FftCap complex entries (28672 B each: 3584 FP32 or 1792 FP64 entries) and a
twiddle table of TW_CAP = 3 x 64 entries.  A chunk of L lines of length N
(L = max(1, min(FftCap / N, lines of the axis))) is loaded, passed through the
out-of-place Stockham passes of N's radix schedule (4, 2, 3, 5, 7 in that
order) and stored.  For every grid edge the batch path admits (at most 64,
factors 2, 3, 5, 7) and every number of lines an axis can have (the product
of two admitted edges), this walks every chunk and every butterfly of every
pass and asserts that each LDS read, write and twiddle index is inside its
buffer: the index arithmetic itself cannot run past the buffers, for any
cell shape and for any of the tuner's workgroup sizes (they change only the
thread stride of the same index set).
"""
import itertools

import numpy as np


def radices(n):
    """The radix schedule of nb_factor: fours, then twos, threes, fives, sevens;
    None when n has another prime factor."""
    r = []
    for p in (4, 2, 3, 5, 7):
        while n % p == 0:
            r.append(p)
            n //= p
    return r if n == 1 else None


ADMITTED = [k for k in range(2, 65) if radices(k) is not None]
FFT_CAP = {"float": 28672 // 8, "double": 28672 // 16}
TW_CAP = 3 * 64


def pass_indices(N, nlines, line_fast, R, Ns):
    """Every (read, write, twiddle) index of one pp_pass, vectorised over the
    butterflies b < nlines * N / R (sl = N, se = 1 for rows; sl = 1, se =
    nlines for interleaved lines)."""
    nb = N // R
    stride = N // (Ns * R)
    sl, se = (1, nlines) if line_fast else (N, 1)
    b = np.arange(nlines * nb)
    if line_fast:
        line, j = b % nlines, b // nlines
    else:
        j, line = b % nb, b // nb
    jm = j % Ns
    k = np.arange(R)[:, None]
    reads = line * sl + (j + k * nb) * se
    writes = line * sl + ((j // Ns) * Ns * R + jm + k * Ns) * se
    tw = jm * k * stride
    return reads, writes, tw


def test_radix_schedules_match_the_admitted_edges():
    assert 64 in ADMITTED and 49 in ADMITTED and 11 not in ADMITTED and 22 not in ADMITTED
    for n in ADMITTED:
        assert int(np.prod(radices(n))) == n


def test_every_lds_index_of_the_batch_transforms_is_in_bounds():
    products = sorted({a * b for a, b in itertools.combinations_with_replacement(ADMITTED, 2)})
    for real, cap in FFT_CAP.items():
        for N in ADMITTED:
            for P in products:
                L = max(1, min(cap // N, P))
                # a chunk of cnt <= L lines: load / store indices t < cnt * N
                assert L * N <= cap, (real, N, P)
                for cnt in {L, P % L} - {0}:
                    for line_fast in (False, True):
                        Ns = 1
                        for R in radices(N):
                            rd, wr, tw = pass_indices(N, cnt, line_fast, R, Ns)
                            assert rd.min() >= 0 and rd.max() < cnt * N <= cap, (real, N, P, cnt, R, Ns)
                            assert wr.min() >= 0 and wr.max() < cnt * N <= cap, (real, N, P, cnt, R, Ns)
                            # every element of the chunk is written exactly once per pass
                            assert np.unique(wr).size == cnt * N, (real, N, P, cnt, R, Ns)
                            assert tw.max() < N <= 64, (N, R, Ns)
                            Ns *= R


def test_twiddle_tables_of_three_axes_fit():
    # the three axes' tables are stored one after the other: K0 + K1 + K2 <= 3 x 64
    assert 3 * max(ADMITTED) <= TW_CAP
