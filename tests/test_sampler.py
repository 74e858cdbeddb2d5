"""The realization kernels' Poisson sampler: CDF windows built on the GPU against scipy's Poisson
CDF, single inversions from given words against exact integer arithmetic, and the xoshiro128**
generator (GF(2)-linear state transition of full period 2^128 - 1, uniform outputs)."""
import numpy as np
import pytest

# 2^128 - 1 = F0 F1 F2 F3 F4 F5 F6 (Fermat numbers), F5 = 641 * 6700417, F6 = 274177 * 67280421310721
PERIOD = (1 << 128) - 1
PERIOD_PRIMES = (3, 5, 17, 257, 65537, 641, 6700417, 274177, 67280421310721)
TAIL = 2.0**-56   # each side of a window drops less than this


def _bits(state):
    return np.array([(state[i // 32] >> (i % 32)) & 1 for i in range(128)], dtype=np.int64)


def _step_matrix():
    from holodeck_amd import _lib
    M = np.zeros((128, 128), dtype=np.int64)
    for i in range(128):
        e = [0, 0, 0, 0]
        e[i // 32] = 1 << (i % 32)
        M[:, i] = _bits(_lib.xoshiro128ss(e, 1)[0])
    return M


def _gf2_pow(M, e):
    R = np.eye(128, dtype=np.int64)
    B = M.copy()
    while e:
        if e & 1:
            R = (R @ B) & 1
        B = (B @ B) & 1
        e >>= 1
    return R


def test_period_primes_multiply_to_the_period():
    assert np.prod([int(p) for p in PERIOD_PRIMES], dtype=object) == PERIOD


def test_xoshiro_state_transition_is_linear_with_full_period():
    from holodeck_amd import _lib
    M = _step_matrix()
    rng = np.random.default_rng(3)
    for _ in range(8):   # the step is the linear map M (no carries anywhere)
        s = [int(v) for v in rng.integers(0, 1 << 32, 4, dtype=np.uint64)]
        assert np.array_equal(_bits(_lib.xoshiro128ss(s, 1)[0]), (M @ _bits(s)) & 1)
    I = np.eye(128, dtype=np.int64)
    assert np.array_equal(_gf2_pow(M, PERIOD), I)
    for p in PERIOD_PRIMES:   # no smaller period: the order of M is exactly 2^128 - 1
        assert not np.array_equal(_gf2_pow(M, PERIOD // p), I), p
    s = [1, 2, 3, 4]   # n steps at once equal n single steps
    st, out = _lib.xoshiro128ss(s, 5)
    t = s
    for k in range(5):
        t, o = _lib.xoshiro128ss(t, 1)
        assert o[0] == out[k]
    assert t == st


def test_xoshiro_outputs_are_uniform():
    from scipy.stats import chisquare
    from holodeck_amd import _lib
    _, w = _lib.xoshiro128ss([0x12345678, 0x9abcdef0, 0x0fedcba9, 0x87654321], 1 << 20)
    for shift in (0, 8, 24):
        counts = np.bincount((w >> shift) & 0xFF, minlength=256)
        assert chisquare(counts).pvalue > 1e-4, shift
    u = w.astype(np.float64) / 2.0**32   # consecutive outputs uncorrelated
    assert abs(np.corrcoef(u[:-1], u[1:])[0, 1]) < 5.0 / np.sqrt(u.size)


def _table(lam, cap=6392):
    """(kb, kt, n, F) with F[j] = floor(CDF(kt + j) 2^64) as Python ints (the CDF summed from the
    support's first count kb), from the GPU build."""
    import torch
    from holodeck_amd import _lib
    from holodeck_amd import device as D
    kn = torch.zeros(3, dtype=torch.int32, device="cuda")
    cdf = torch.zeros(cap, dtype=torch.int64, device="cuda")
    _lib.call("hd_poisson_cdf_table", float(lam), cap, D.ptr(kn), D.ptr(cdf), D.stream_ptr())
    kb, kt, n = (int(v) for v in kn.cpu())
    F = cdf[:n].cpu().numpy().view(np.uint64)
    return kb, kt, n, [int(v) for v in F]


@pytest.mark.gpu
@pytest.mark.parametrize("lam", [1.0, 1.7, 6.3, 30.0, 143.9, 1000.5, 2.5e4, 1.2e5])
def test_cdf_window_matches_the_poisson_cdf(gpu, lam):
    """The table covers lam -/+ 4 sqrt(lam) (clipped to the support), its first entry holds the
    whole CDF below (summed from kb, below which less than 2^-56 is dropped)."""
    from scipy.stats import poisson
    print(f"[cdf table lam={lam}]", flush=True)
    kb, kt, n, F = _table(lam)
    assert n > 0 and kt >= kb
    kc = kt + n - 1
    sd = np.sqrt(lam)
    assert poisson.cdf(kb - 1, lam) < TAIL
    assert kt == kb or kt + 1 <= lam - 4.0 * sd
    assert kc >= lam + 4.0 * sd + 1.0 or poisson.sf(kc, lam) < TAIL
    assert all(a <= b for a, b in zip(F[:-1], F[1:]))   # non-decreasing: inversion is monotone in u
    got = np.array([f / 2.0**64 for f in F])
    expect = poisson.cdf(kt + np.arange(n), lam) - poisson.cdf(kb - 1, lam)
    err = np.abs(got - expect)
    # CDF to ~(kt + n - kb) ulp, then rounded down to a multiple of 2^-64
    assert np.all(err <= 1e-11 * expect + 2.0**-63), (err / expect).max()


def _tail_count(u, lam, kb, kt, F, dps=60):
    """The count the tail walk must give for u / 2^64 outside the table F (entries kt .. kt + n - 1),
    with the table's edge value as the CDF there (the sampler's distribution is the table's) and the
    pmf beyond it in mpmath at dps digits: downwards to kb, upwards to the support's end (Bernstein
    bound, as the kernels)."""
    import mpmath
    n = len(F)
    with mpmath.workdps(dps):
        L = mpmath.mpf(lam)
        two64 = mpmath.mpf(2) ** 64

        def pmf(k):
            return two64 * mpmath.exp(-L + k * mpmath.log(L) - mpmath.loggamma(k + 1)) if k > 0 else two64 * mpmath.exp(-L)

        if u < F[0]:                 # CDF(kt) > u: walk down while CDF(k - 1) > u
            k, c = kt, mpmath.mpf(F[0])
            while k > kb and c - pmf(k) > u:
                c -= pmf(k)
                k -= 1
            return k
        b = 56 * mpmath.log(2)
        khi = int(mpmath.ceil(L + b / 3 + mpmath.sqrt(b * b / 9 + 2 * b * L)))
        k, c = kt + n - 1, mpmath.mpf(F[-1])
        if k >= khi:
            return k
        while True:                  # CDF(k) <= u: the first k with CDF(k) > u
            k += 1
            c += pmf(k)
            if c > u or k >= khi:
                return k


def _dev_tab(F):
    import torch
    t = np.array([f >> 32 for f in F] + [f & 0xFFFFFFFF for f in F], dtype=np.uint64).astype(np.uint32)
    return torch.from_numpy(t.view(np.int32)).cuda()


def _draw_words(lam, kb, kt, F, xs, ws):
    import torch
    from holodeck_amd import _lib
    from holodeck_amd import device as D
    m = len(xs)
    words = np.empty(2 * m, dtype=np.uint32)
    words[0::2] = np.array(xs, dtype=np.uint64).astype(np.uint32)
    words[1::2] = np.array(ws, dtype=np.uint64).astype(np.uint32)
    w_d = torch.from_numpy(words.view(np.int32)).cuda()
    tab_d = _dev_tab(F)
    count = torch.zeros(m, dtype=torch.int32, device="cuda")
    used = torch.zeros(m, dtype=torch.int32, device="cuda")
    _lib.call("hd_table_draw_words", D.ptr(tab_d), len(F), float(lam), kb, kt, D.ptr(w_d), m, D.ptr(count),
              D.ptr(used), D.stream_ptr())
    return count.cpu().numpy(), used.cpu().numpy()


def _table_index(u, F):
    """First entry j with F[j] > u (len(F) if none)."""
    for j, f in enumerate(F):
        if f > u:
            return j
    return len(F)


@pytest.mark.gpu
@pytest.mark.parametrize("lam", [2.2, 37.0, 1000.5])
def test_inversion_from_given_words_matches_exact_arithmetic(gpu, lam):
    """Draws from given words: inside the table exact integer comparisons with its entries; below
    and above it (the tail walks) the count from the table's edge value and the mpmath pmf beyond."""
    print(f"[inversion from words lam={lam}]", flush=True)
    kb, kt, n, F = _table(lam)
    rng = np.random.default_rng(11)
    xs, ws = list(rng.integers(0, 1 << 32, 3000)), list(rng.integers(0, 1 << 32, 3000))
    # first words equal to a table value's high word (and its neighbours): the second word decides
    for f in F[:: max(1, n // 60)] + [F[0], F[-1]]:
        for d in (-1, 0, 1):
            x = (f >> 32) + d
            if 0 <= x < (1 << 32):
                for w in (0, (f & 0xFFFFFFFF), (f & 0xFFFFFFFF) + 1, (1 << 32) - 1, int(rng.integers(0, 1 << 32))):
                    if w < (1 << 32):
                        xs.append(x)
                        ws.append(w)
    # uniforms in the tails: below the first entry (when the lower tail is summed) and above the last
    ntail = 400
    if kt > kb:
        top = F[0] >> 32
        xs += list(rng.integers(0, top + 1, ntail))
        ws += list(rng.integers(0, 1 << 32, ntail))
    xs += list(rng.integers(F[-1] >> 32, 1 << 32, ntail))
    ws += list(rng.integers(0, 1 << 32, ntail))
    xs.append((1 << 32) - 1)
    ws.append((1 << 32) - 1)
    count, used = _draw_words(lam, kb, kt, F, xs, ws)
    n_tail = 0
    for i, (x, w) in enumerate(zip(xs, ws)):
        u = (int(x) << 32) | int(w)
        j = _table_index(u, F)
        if (j == 0 and kt > kb) or j == n:
            want = _tail_count(u, lam, kb, kt, F)
            n_tail += 1
            assert used[i] == 2, i
        else:
            want = kt + j
        assert count[i] == want, (i, x, w, count[i], want)
    assert n_tail >= ntail
    assert set(np.unique(used)) <= {1, 2} and (used == 2).sum() >= n
    one = np.flatnonzero(used == 1)   # decided by the first word: the second cannot matter
    for i in one[::37]:
        lo_u, hi_u = int(xs[i]) << 32, (int(xs[i]) << 32) | 0xFFFFFFFF
        assert kt + _table_index(lo_u, F) == kt + _table_index(hi_u, F) == count[i]


def _opens(x, F):
    """True when the first table value with high word >= x has high word x (the low words decide)."""
    for f in F:
        if (f >> 32) >= x:
            return (f >> 32) == x
    return False


@pytest.mark.gpu
def test_paired_inversions_take_words_in_order(gpu):
    """The kernels draw records in pairs with interleaved searches: words x_a, x_b, then w_a, w_b
    where a tie needs them, then where a tail walk needs them."""
    import torch
    from holodeck_amd import _lib
    from holodeck_amd import device as D
    print("[paired inversions]", flush=True)
    la, lb = 2.2, 37.0
    kba, kta, na, Fa = _table(la)
    kbb, ktb, nb, Fb = _table(lb)
    ta, tb = _dev_tab(Fa), _dev_tab(Fb)
    rng = np.random.default_rng(5)
    quads = [list(rng.integers(0, 1 << 32, 4)) for _ in range(2000)]
    for fa, fb in zip(Fa, Fb[::3]):   # both, one or neither open
        for xa, xb in ((fa >> 32, fb >> 32), (fa >> 32, int(rng.integers(0, 1 << 32))),
                       (int(rng.integers(0, 1 << 32)), fb >> 32)):
            quads.append([xa, xb, int(rng.integers(0, 1 << 32)), int(rng.integers(0, 1 << 32))])
    for _ in range(200):               # tails of either or both
        xa = int(rng.integers(Fa[-1] >> 32, 1 << 32)) if rng.random() < 0.5 else int(rng.integers(0, 1 << 32))
        xb = int(rng.integers(0, (Fb[0] >> 32) + 1)) if rng.random() < 0.5 else int(rng.integers(Fb[-1] >> 32, 1 << 32))
        quads.append([xa, xb, int(rng.integers(0, 1 << 32)), int(rng.integers(0, 1 << 32))])
    m = len(quads)
    words = np.array(quads, dtype=np.uint64).astype(np.uint32).ravel()
    w_d = torch.from_numpy(words.view(np.int32)).cuda()
    count = torch.zeros(2 * m, dtype=torch.int32, device="cuda")
    used = torch.zeros(m, dtype=torch.int32, device="cuda")
    _lib.call("hd_table_draw_pair_words", D.ptr(ta), na, la, kba, kta, D.ptr(tb), nb, lb, kbb, ktb, D.ptr(w_d), m,
              D.ptr(count), D.ptr(used), D.stream_ptr())
    count, used = count.cpu().numpy().reshape(m, 2), used.cpu().numpy()
    n_open = n_tail = 0
    for i, quad in enumerate(quads):
        xa, xb, w1, w2 = (int(v) for v in quad)
        rest = [w1, w2]
        oa, ob = _opens(xa, Fa), _opens(xb, Fb)
        wa = rest.pop(0) if oa else None
        wb = rest.pop(0) if ob else None

        def result(x, w, kb, kt, F):
            # not open: the first word alone decides the entry (any second word gives the same)
            j = _table_index((x << 32) | (w if w is not None else 0), F)
            return j, (j == 0 and kt > kb) or j == len(F)

        ja, tail_a = result(xa, wa, kba, kta, Fa)
        jb, tail_b = result(xb, wb, kbb, ktb, Fb)
        if tail_a and wa is None:
            wa = rest.pop(0)
        if tail_b and wb is None:
            wb = rest.pop(0)
        want_a = _tail_count((xa << 32) | wa, la, kba, kta, Fa) if tail_a else kta + ja
        want_b = _tail_count((xb << 32) | wb, lb, kbb, ktb, Fb) if tail_b else ktb + jb
        assert count[i, 0] == want_a and count[i, 1] == want_b, (i, count[i], want_a, want_b)
        assert used[i] == 2 + (wa is not None) + (wb is not None), i
        n_open += oa + ob
        n_tail += tail_a + tail_b
    assert n_open > na and n_tail >= 100
