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
    """(klo, n, F) with F[k] = floor(CDF(klo + k) 2^64) as Python ints, from the GPU build."""
    import torch
    from holodeck_amd import _lib
    from holodeck_amd import device as D
    kn = torch.zeros(2, dtype=torch.int32, device="cuda")
    cdf = torch.zeros(cap, dtype=torch.int64, device="cuda")
    _lib.call("hd_poisson_cdf_table", float(lam), cap, D.ptr(kn), D.ptr(cdf), D.stream_ptr())
    klo, n = (int(v) for v in kn.cpu())
    F = cdf[:n].cpu().numpy().view(np.uint64)
    return klo, n, [int(v) for v in F]


@pytest.mark.gpu
@pytest.mark.parametrize("lam", [1.0, 1.7, 6.3, 30.0, 143.9, 1000.5, 2.5e4, 1.2e5])
def test_cdf_window_matches_the_poisson_cdf(gpu, lam):
    from scipy.stats import poisson
    print(f"[cdf window lam={lam}]", flush=True)
    klo, n, F = _table(lam)
    assert n > 0
    assert poisson.cdf(klo - 1, lam) < TAIL and poisson.sf(klo + n - 1, lam) < TAIL
    assert all(a <= b for a, b in zip(F[:-1], F[1:]))   # non-decreasing: inversion is monotone in u
    got = np.array([f / 2.0**64 for f in F])
    expect = poisson.cdf(klo + np.arange(n), lam) - poisson.cdf(klo - 1, lam)
    err = np.abs(got - expect)
    # CDF to ~n ulp, then rounded down to a multiple of 2^-64
    assert np.all(err <= 1e-11 * expect + 2.0**-63), (err / expect).max()
    assert abs(got[-1] - 1.0) < 1e-12


def _exact_index(x, w, F):
    """Smallest k with F[k] > u64 = x 2^32 + w (u = u64 / 2^64); n - 1 if none."""
    u = (int(x) << 32) | int(w)
    for k, f in enumerate(F):
        if f > u:
            return k
    return len(F) - 1


@pytest.mark.gpu
@pytest.mark.parametrize("lam", [2.2, 37.0])
def test_inversion_from_given_words_matches_exact_arithmetic(gpu, lam):
    import torch
    from holodeck_amd import _lib
    from holodeck_amd import device as D
    print(f"[inversion from words lam={lam}]", flush=True)
    klo, n, F = _table(lam)
    tab = np.array([f >> 32 for f in F] + [f & 0xFFFFFFFF for f in F], dtype=np.uint64).astype(np.uint32)
    tab_d = torch.from_numpy(tab.view(np.int32)).cuda()
    rng = np.random.default_rng(11)
    xs, ws = list(rng.integers(0, 1 << 32, 3000)), list(rng.integers(0, 1 << 32, 3000))
    # first words equal to a table value's high word (and its neighbours): the second word decides
    for f in F:
        for d in (-1, 0, 1):
            x = (f >> 32) + d
            if 0 <= x < (1 << 32):
                for w in (0, (f & 0xFFFFFFFF), (f & 0xFFFFFFFF) + 1, (1 << 32) - 1, int(rng.integers(0, 1 << 32))):
                    if w < (1 << 32):
                        xs.append(x)
                        ws.append(w)
    xs.append((1 << 32) - 1)
    ws.append((1 << 32) - 1)
    m = len(xs)
    words = np.empty(2 * m, dtype=np.uint32)
    words[0::2] = np.array(xs, dtype=np.uint64).astype(np.uint32)
    words[1::2] = np.array(ws, dtype=np.uint64).astype(np.uint32)
    w_d = torch.from_numpy(words.view(np.int32)).cuda()
    slot = torch.zeros(m, dtype=torch.int32, device="cuda")
    used = torch.zeros(m, dtype=torch.int32, device="cuda")
    _lib.call("hd_table_draw_words", D.ptr(tab_d), n, D.ptr(w_d), m, D.ptr(slot), D.ptr(used), D.stream_ptr())
    slot, used = slot.cpu().numpy(), used.cpu().numpy()
    expect = np.array([_exact_index(x, w, F) for x, w in zip(xs, ws)])
    assert np.array_equal(slot, expect), np.flatnonzero(slot != expect)[:10]
    assert set(np.unique(used)) <= {1, 2} and (used == 2).sum() >= n
    one = np.flatnonzero(used == 1)   # decided by the first word: the second cannot matter
    for i in one[::37]:
        assert _exact_index(xs[i], 0, F) == _exact_index(xs[i], (1 << 32) - 1, F) == slot[i]


def _opens(x, F):
    """True when the first table value with high word >= x has high word x (the low words decide)."""
    for f in F:
        if (f >> 32) >= x:
            return (f >> 32) == x
    return False


@pytest.mark.gpu
def test_paired_inversions_take_words_in_order(gpu):
    """The kernels draw records in pairs with interleaved searches: words x_a, x_b, then w_a, w_b."""
    import torch
    from holodeck_amd import _lib
    from holodeck_amd import device as D
    print("[paired inversions]", flush=True)
    _, na, Fa = _table(2.2)
    _, nb, Fb = _table(37.0)

    def dev_tab(F):
        t = np.array([f >> 32 for f in F] + [f & 0xFFFFFFFF for f in F], dtype=np.uint64).astype(np.uint32)
        return torch.from_numpy(t.view(np.int32)).cuda()

    ta, tb = dev_tab(Fa), dev_tab(Fb)
    rng = np.random.default_rng(5)
    quads = [list(rng.integers(0, 1 << 32, 4)) for _ in range(2000)]
    for fa, fb in zip(Fa, Fb[::3]):   # both, one or neither open
        for xa, xb in ((fa >> 32, fb >> 32), (fa >> 32, int(rng.integers(0, 1 << 32))),
                       (int(rng.integers(0, 1 << 32)), fb >> 32)):
            quads.append([xa, xb, int(rng.integers(0, 1 << 32)), int(rng.integers(0, 1 << 32))])
    m = len(quads)
    words = np.array(quads, dtype=np.uint64).astype(np.uint32).ravel()
    w_d = torch.from_numpy(words.view(np.int32)).cuda()
    slot = torch.zeros(2 * m, dtype=torch.int32, device="cuda")
    used = torch.zeros(m, dtype=torch.int32, device="cuda")
    _lib.call("hd_table_draw_pair_words", D.ptr(ta), na, D.ptr(tb), nb, D.ptr(w_d), m, D.ptr(slot), D.ptr(used),
              D.stream_ptr())
    slot, used = slot.cpu().numpy().reshape(m, 2), used.cpu().numpy()
    n_open = 0
    for i, (xa, xb, w1, w2) in enumerate(quads):
        oa, ob = _opens(xa, Fa), _opens(xb, Fb)
        wa = w1 if oa else 0
        wb = (w2 if oa else w1) if ob else 0
        assert slot[i, 0] == _exact_index(xa, wa, Fa) and slot[i, 1] == _exact_index(xb, wb, Fb), i
        assert used[i] == 2 + oa + ob, i
        n_open += oa + ob
    assert n_open > na
