"""Host-side checks of the LDS layouts of the radix-8 transforms in
rot2d3c_amd/csrc/fft_core.h (Stockham passes, and the in-place form used at
n = 2048 and 8192).  The index functions are restated from that header; the
checks are properties the kernels rely on and that a wrong digit order, skew
or size would break:
- the in-place first pass writes every position exactly once, and the slot
  map is injective and stays inside the LDS array the kernels allocate
  (Shape<N>::LDSN), so no access can leave it;
- each middle pass's butterflies own disjoint position sets (writing back in
  place needs no barrier);
- the in-place and Stockham layouts feed every butterfly the same partial
  results, so both give the same bits, and both equal the DFT;
- with 8 lanes per 128-byte LDS cycle, the in-place accesses are free of bank
  conflicts at n = 2048 and 8192 (the sizes that use it).
No GPU is needed."""
import numpy as np
import pytest


def shape(n):
    log = n.bit_length() - 1
    nmid = (log - 4) // 3
    rl = 1 << (log - 3 - 3 * nmid)
    return n // 8, log, nmid, rl


def pad(i):  # one pad slot per 16 entries (PS = 4)
    return i + (i >> 4)


def ip_rho(n, b):
    t, log, nmid, rl = shape(n)
    rho, f = 0, 1
    for s in range(2, nmid + 2):
        rho += ((b >> (log - 3 * s)) & 7) * f
        f *= 8
    return rho + (b & (rl - 1)) * f


def ip_slot(n, p):
    t, log, nmid, rl = shape(n)
    sk2 = 2 if rl == 2 else 0
    return pad(p) + p // (n // rl) + sk2 * (p // (n // 16))


def ldsn_ip(n):
    t, log, nmid, rl = shape(n)
    sk2 = 2 if rl == 2 else 0
    return n + (n >> 4) + rl + 15 * sk2


def middle_positions(n, L, b):
    j = b & (L - 1)
    return [(b - j) * 8 + j + m * L for m in range(8)]


@pytest.mark.parametrize("n", [64, 128, 256, 512, 1024, 2048, 4096, 8192])
def test_inplace_positions_and_slots(n):
    t, log, nmid, rl = shape(n)
    first = sorted(m + 8 * ip_rho(n, b) for b in range(t) for m in range(8))
    assert first == list(range(n))
    slots = [ip_slot(n, p) for p in range(n)]
    assert len(set(slots)) == n and min(slots) >= 0 and max(slots) < ldsn_ip(n)
    L = 8
    for _ in range(nmid):
        pos = [p for b in range(t) for p in middle_positions(n, L, b)]
        assert sorted(pos) == list(range(n))
        L *= 8
    d = n // rl
    last = sorted(j + m * d for b in range(t) for u in range(8 // rl) for j in [b + t * u]
                  for m in range(rl))
    assert last == list(range(n))


def _dft(x, s):
    r = len(x)
    k = np.arange(r)
    return np.exp(s * 2j * np.pi * np.outer(k, k) / r) @ np.asarray(x)


def _transform(n, s, x, inplace):
    """Index-level emulation of fft_canon: the same butterflies and twiddles,
    with the Stockham or the in-place LDS positions."""
    t, log, nmid, rl = shape(n)
    d = n // rl
    tw = np.exp(-2j * np.pi * np.arange(n) / n)
    tws = tw if s < 0 else np.conj(tw)
    lds = np.full(ldsn_ip(n), np.nan, dtype=complex)
    slot = (lambda p: ip_slot(n, p)) if inplace else pad
    for b in range(t):
        y = _dft(x[b + t * np.arange(8)], s)
        for m in range(8):
            lds[slot(m + 8 * ip_rho(n, b) if inplace else 8 * b + m)] = y[m]
    L = 8
    for _ in range(nmid):
        new = lds.copy()
        for b in range(t):
            j = b & (L - 1)
            st = (n // (8 * L)) * j
            if inplace:
                src = dst = middle_positions(n, L, b)
            else:
                src = [b + m * t for m in range(8)]
                dst = middle_positions(n, L, b)
            xs = [lds[slot(p)] * tws[st * m % n] for m, p in enumerate(src)]
            for p, v in zip(dst, _dft(xs, s)):
                new[slot(p)] = v
        lds = new
        L *= 8
    out = np.zeros(n, dtype=complex)
    for b in range(t):
        for u in range(8 // rl):
            j = b + t * u
            xs = [lds[slot(j + m * d)] * tws[j * m % n] for m in range(rl)]
            out[j + d * np.arange(rl)] = _dft(xs, s)
    return out


@pytest.mark.parametrize("n", [64, 128, 256, 512, 2048])
def test_inplace_equals_stockham_and_dft(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    for s in (-1, 1):
        a, b = _transform(n, s, x, False), _transform(n, s, x, True)
        assert np.array_equal(a, b)
        ref = np.fft.fft(x) if s < 0 else np.fft.ifft(x) * n
        assert np.abs(b - ref).max() <= 1e-12 * np.abs(ref).max()


def _worst_conflict(slots):
    """Largest number of lanes of one 8-lane group that hit the same 16-byte
    bank group (slot mod 8)."""
    worst = 1
    for g in range(0, len(slots), 8):
        worst = max(worst, int(np.bincount(np.asarray(slots[g:g + 8]) % 8).max()))
    return worst


@pytest.mark.parametrize("n,il", [(2048, 1), (2048, 2), (8192, 1)])
def test_inplace_is_bank_conflict_free(n, il):
    t, log, nmid, rl = shape(n)
    lanes = np.arange(t * il)
    r, b = lanes % il, lanes // il

    def worst(pos):
        return _worst_conflict([ip_slot(n, p) * il + rr for p, rr in zip(pos, r)])

    for m in range(8):
        assert worst([m + 8 * ip_rho(n, bb) for bb in b]) == 1
    L = 8
    for _ in range(nmid):
        for m in range(8):
            assert worst([middle_positions(n, L, bb)[m] for bb in b]) == 1
        L *= 8
    for u in range(8 // rl):
        for m in range(rl):
            assert worst([bb + t * u + m * (n // rl) for bb in b]) == 1
