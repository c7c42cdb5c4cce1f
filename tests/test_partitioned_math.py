"""CPU pin of the partitioned line-solve algorithm that the HIP kernel
k_cn_part implements (segment elimination of Laszlo, Giles & Appleyard,
ACM TOMS 42(4) 2016, then Hockney's parallel cyclic reduction of the
segment ends), restated here in numpy step for step, with the kernel's
padding (identity rows past n) and its lane layout (LP segments of M rows,
two reduced equations per lane, level 0 pairing 2p with 2p-1 / 2p+1).

It is checked against scipy's banded direct solve on theta-scheme matrices
of the reference's form (src/Diffusion.F90:807-853): random row factors
over five decades, strongly varying cell volumes (where dgtsv swaps rows),
theta 0, 0.3, 0.5 and 1, every (LP, M) shape the kernel uses and line
lengths 3..LP*M.  A dropped term, a wrong sign or a shifted neighbour in
any of the recurrences fails this by orders of magnitude.
"""
import numpy as np
import pytest
from scipy.linalg import solve_banded


def cn_matrix(L, R, th):
    n = len(L)
    a = -(th * L)
    c = -(th * R)
    a[0] = 0.0
    c[-1] = 0.0
    b = 1.0 + th * (L + R)
    b[0] -= th * L[0]
    b[-1] -= th * R[-1]
    return a, b, c


def part_solve(a, b, c, d, LP, M, mistake=None):
    n = len(a)
    assert n <= LP * M
    npad = LP * M
    A = np.zeros(npad)
    B = np.ones(npad)
    C = np.zeros(npad)
    D = np.zeros(npad)
    A[:n], B[:n], C[:n], D[:n] = a, b, c, d
    ah = np.zeros((LP, M))
    ch = np.zeros((LP, M))
    dh = np.zeros((LP, M))
    for p in range(LP):
        s = p * M
        al = np.zeros(M)
        be = np.zeros(M)
        for k in range(M):
            ak, bk, ck = A[s + k], B[s + k], C[s + k]
            if k < 2:
                al[k] = 1.0 / bk
                ah[p, k] = ak * al[k]
                ch[p, k] = ck * al[k]
            else:
                al[k] = 1.0 / (bk - ak * ch[p, k - 1])
                be[k] = al[k] * ak
                ah[p, k] = -be[k] * ah[p, k - 1]
                ch[p, k] = al[k] * ck
        ga = ch[p].copy()
        for k in range(M - 3, 0, -1):
            ah[p, k] = ah[p, k] - ga[k] * ah[p, k + 1]
            ch[p, k] = -ga[k] * ch[p, k + 1]
        rho = 1.0 / (1.0 - ga[0] * ah[p, 1])
        ah[p, 0] = rho * ah[p, 0]
        ch[p, 0] = -rho * ga[0] * ch[p, 1]
        # right-hand side
        for k in range(M):
            dh[p, k] = al[k] * D[s + k] - (be[k] * dh[p, k - 1] if k >= 2 else 0.0)
        for k in range(M - 3, 0, -1):
            dh[p, k] -= (-1.0 if mistake == "bwd_sign" else 1.0) * ga[k] * dh[p, k + 1]
        dh[p, 0] = rho * (dh[p, 0] - ga[0] * dh[p, 1])
    # reduced system: equation 2p = first row of segment p, 2p+1 = its last
    RA = np.stack([ah[:, 0], ah[:, M - 1]], axis=1)
    RB = np.ones((LP, 2))
    RC = np.stack([ch[:, 0], ch[:, M - 1]], axis=1)
    RD = np.stack([dh[:, 0], dh[:, M - 1]], axis=1)

    def nb(X, q, j, ident):
        return X[q, j] if 0 <= q < LP else ident

    lev = 0
    while (1 << lev) < 2 * LP:
        nA, nB, nC, nD = RA.copy(), RB.copy(), RC.copy(), RD.copy()
        for p in range(LP):
            for j in range(2):
                if lev == 0:
                    mq, mj = (p - 1, 1) if j == 0 else (p, 0)
                    pq, pj = (p, 1) if j == 0 else (p + 1, 0)
                else:
                    h = 1 << (lev - 1)
                    if mistake == "pcr_stride":
                        h = max(1, h // 2)
                    mq, mj, pq, pj = p - h, j, p + h, j
                k1 = RA[p, j] / nb(RB, mq, mj, 1.0)
                k2 = RC[p, j] / nb(RB, pq, pj, 1.0)
                nA[p, j] = -k1 * nb(RA, mq, mj, 0.0)
                nC[p, j] = -k2 * nb(RC, pq, pj, 0.0)
                nB[p, j] = RB[p, j] - k1 * nb(RC, mq, mj, 0.0) - k2 * nb(RA, pq, pj, 0.0)
                nD[p, j] = RD[p, j] - k1 * nb(RD, mq, mj, 0.0) - k2 * nb(RD, pq, pj, 0.0)
        RA, RB, RC, RD = nA, nB, nC, nD
        lev += 1
    z = RD / RB
    x = np.zeros(npad)
    for p in range(LP):
        F, E = z[p, 0], z[p, 1]
        x[p * M] = F
        x[p * M + M - 1] = E
        for k in range(1, M - 1):
            x[p * M + k] = dh[p, k] - ah[p, k] * (E if mistake == "swap_ends" else F) - ch[p, k] * E
    return x[:n]


# the (lanes per line, rows per lane) shapes k_cn_part is launched with
SHAPES = [(64, 4), (64, 6), (64, 8), (16, 4), (32, 4)]


@pytest.mark.parametrize("LP,M", SHAPES)
@pytest.mark.parametrize("th", [0.0, 0.3, 0.5, 1.0])
def test_partitioned_solve_matches_direct_solve(LP, M, th):
    rng = np.random.default_rng(LP * 100 + M + int(10 * th))
    worst = 0.0
    for trial in range(12):
        n = int(rng.integers(3, LP * M + 1)) if trial else LP * M
        L = rng.random(n) * 10 ** rng.uniform(-2, 3)
        R = rng.random(n) * 10 ** rng.uniform(-2, 3)
        if trial % 3 == 0:
            L *= 10 ** rng.uniform(-3, 3, n)  # neighbouring volumes differ a lot
        a, b, c = cn_matrix(L, R, th)
        d = rng.standard_normal(n)
        x = part_solve(a, b, c, d, LP, M)
        ab = np.zeros((3, n))
        ab[0, 1:] = c[:-1]
        ab[1] = b
        ab[2, :-1] = a[1:]
        xr = solve_banded((1, 1), ab, d)
        worst = max(worst, np.max(np.abs(x - xr)) / np.max(np.abs(xr)))
    assert worst < 1e-12, worst


@pytest.mark.parametrize("mistake", ["bwd_sign", "pcr_stride", "swap_ends"])
def test_a_planted_mistake_is_caught(mistake):
    """The same check with a mistake planted in the algorithm (a sign in the
    backward elimination, the PCR neighbour distance, the end value used for
    the interior rows) fails: the pin is sensitive to the recurrences."""
    rng = np.random.default_rng(5)
    n, LP, M = 200, 64, 4
    L = rng.random(n) * 50
    R = rng.random(n) * 50
    a, b, c = cn_matrix(L, R, 0.5)
    d = rng.standard_normal(n)
    ab = np.zeros((3, n))
    ab[0, 1:] = c[:-1]
    ab[1] = b
    ab[2, :-1] = a[1:]
    xr = solve_banded((1, 1), ab, d)
    good = part_solve(a, b, c, d, LP, M)
    assert np.max(np.abs(good - xr)) / np.max(np.abs(xr)) < 1e-12
    bad = part_solve(a, b, c, d, LP, M, mistake=mistake)
    assert np.max(np.abs(bad - xr)) / np.max(np.abs(xr)) > 1e-6
