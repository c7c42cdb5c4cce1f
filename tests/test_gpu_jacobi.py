"""Jacobi sweeps for the y and z theta-scheme lines that the partitioned
solves leave to dgtsv order (line_solver="partitioned", lines longer than
part_max_line_strided), and for the x lines that seven sweeps solve (the
partitioned solves taking the other x lines), against the oracle's
dgtsv-order solves.

Row t of a line gives x_t = y_t + a_t x_{t-1} + b_t x_{t+1} (y = B / D,
a = th L / D, b = th R / D); K + 1 sweeps from zero leave an error of at
most rho^(K+1) of the line's largest |x|, rho = max (|DL| + |DU|) / |D|
over the sweep's rows.  The library reads rho back from the factorisation
and takes the least K + 1 <= 7 with rho^(K+1) <= 2^-52 (1 - rho), else
dgtsv order.

* rho is recomputed here from the row factors of the state (numpy, the
  reference's formulas for L and R) and the number of sweeps the library
  reports must be the one that rho asks for: the device-side reduction of
  rho and its readback are pinned.
* Results within the 1e-12 relative L-inf bar of the oracle for 1 to 7
  sweeps (x lines: up to 12), theta 0, 0.5 and 1, 1, 2 and 4 bins per thread, subgrid bounds
  and an IsAloft mask; theta = 0 (rho = 0, one sweep: x = B exactly) is
  bitwise.
* One sweep fewer than rho asks for, sweeps switched off, or rho too large
  for 7 sweeps: dgtsv order, bitwise the oracle.
* The first sweep after a factorisation (rho not yet on the host) runs in
  dgtsv order; full steps with glue and a source over several steps stay
  within the bar.
* x lines (one workgroup per line, the rows across its lanes and waves):
  every line, some or none of them needing at most 12 sweeps (per-line
  rho recomputed here; the sweeps those lines take pinned), periodic and not,
  lines of 600 cells over three waves with subgrid bounds and a mask;
  lines of one wave stay in the partitioned solves by default.
* The adjoint (lambda <- K B p(A) K^-1 lambda, the transpose of the
  sweeps' map p(A) B): <F q, w> = <q, F^T w> against the oracle's forward
  map for y and z lines with 2-7 sweeps, theta 0 and 1, subgrids and a
  mask, x lines split with the partitioned adjoint (one wave) and with
  the transposed dgtsv-order adjoint (600 cells).
"""
import math

import numpy as np
import pytest

from test_gpu_parity import ARRAYS, assert_same

pytestmark = pytest.mark.gpu

TOL = 1e-12  # relative L-inf, the parity tolerance for floating-point results
MAX_SWEEPS = 7     # y and z lines
MAX_SWEEPS_X = 12  # x lines (the library's default cap)


@pytest.fixture(scope="module")
def tr(pkg):
    import importlib
    t = importlib.import_module("volcano-ash3d_amd.transport")
    t.load()
    return t


def line_factors(st, dirn):
    """Row factors L, R of every interior cell for lines along dirn (0: x,
    1: y, 2: z), src/Diffusion.F90:807-816: kds_I = 0.5 (K_{I-1} + K_I)
    sig_I / (0.5 (kap_{I-1} + kap_I)), L = dt kds_c sig_c / kap_c, R = dt
    kds_{c+1} sig_{c+1} / kap_c.  Arrays indexed [line a, line b, row]
    (x lines: a = j, b = k; y lines: i, k; z lines: i, j)."""
    nx, ny, nz = st.nx, st.ny, st.nz
    kk = (st.kx, st.ky, st.kz)[dirn]
    sg = (st.sigma_nx_pd, st.sigma_ny_pd, st.sigma_nz_pd)[dirn]
    # cells 0..n+1 in every dimension (kappa and sigma carry two ghosts, K one)
    kap = st.kappa_pd[1:nx + 3, 1:ny + 3, 1:nz + 3]
    sig = sg[1:nx + 3, 1:ny + 3, 1:nz + 3]
    K = kk[0:nx + 2, 0:ny + 2, 0:nz + 2]
    idx = [slice(1, nx + 1), slice(1, ny + 1), slice(1, nz + 1)]
    idx[dirn] = slice(None)
    kap, sig, K = (np.moveaxis(v[tuple(idx)], dirn, -1) for v in (kap, sig, K))
    n = kap.shape[2] - 2
    kds = 0.5 * (K[..., 0:n + 1] + K[..., 1:n + 2]) * (sig[..., 1:n + 2] / (0.5 * (kap[..., 0:n + 1] + kap[..., 1:n + 2])))
    kc = kap[..., 1:n + 1]
    L = st.dt * kds[..., 0:n] * sig[..., 1:n + 1] / kc
    R = st.dt * kds[..., 1:n + 1] * sig[..., 2:n + 2] / kc
    return L, R


def line_rho(st, dirn, rmin=1, rmax=None):
    """max (|DL| + |DU|) / |D| over rows rmin..rmax of every line."""
    return float(per_line_rho(st, dirn, rmin, rmax).max())


def per_line_rho(st, dirn, rmin=1, rmax=None):
    """max (|DL| + |DU|) / |D| over rows rmin..rmax, per line."""
    L, R = line_factors(st, dirn)
    rmax = L.shape[2] if rmax is None else rmax
    L, R = L[..., rmin - 1:rmax], R[..., rmin - 1:rmax]
    th = st.imp_fac
    D = 1.0 + th * (L + R)
    D[..., 0] -= th * L[..., 0]
    D[..., -1] -= th * R[..., -1]
    DL = th * L
    DL[..., 0] = 0.0
    DU = th * R
    DU[..., -1] = 0.0
    return ((np.abs(DL) + np.abs(DU)) / np.abs(D)).max(axis=-1)


def x_split(st, rmin=1, rmax=None, max_sweeps=MAX_SWEEPS_X):
    """(lines taking Jacobi sweeps, all lines, the most sweeps a line takes)
    for x lines: each line the sweeps its own rho needs, if at most
    max_sweeps."""
    sw = np.vectorize(lambda r: expected_sweeps(r, max_sweeps))(per_line_rho(st, 0, rmin, rmax))
    j = sw > 0
    return int(j.sum()), sw.size, int(sw[j].max()) if j.any() else 0


def expected_sweeps(rho, max_sweeps=MAX_SWEEPS):
    """The least K + 1 <= max_sweeps with rho^(K+1) <= 2^-52 (1 - rho), or 0."""
    bound = math.ldexp(1.0, -52) * (1.0 - rho)
    for sweeps in range(1, max_sweeps + 1):
        if rho ** sweeps <= bound:
            return sweeps
    return 0


def run(pkg, oracle, tr, st, calls, after=None, **tune):
    """The same calls on the oracle and the device (synchronised after each
    call, so that the factorisation's rho is on the host for the next);
    after(T, k) runs after device call k."""
    a, b = st.copy(), st.copy()
    O = oracle.Oracle(a)
    T = tr.Transport(b)
    T.tune(line_solver="partitioned", part_max_line_strided=3, **tune)
    for k, (name, *args) in enumerate(calls):
        getattr(O, name)(*args)
        getattr(T, name)(*args)
        T.synchronize()
        if after:
            after(T, k)
    T.sync_to_host()
    T.close()
    return a, b


# dt scale factors of the small state below and the sweeps they need:
# y lines 2..7, z lines 2, 3, 5, 6, 7
SCALES = [(1, 1e-3), (1, 0.03), (1, 1.0), (1, 10.0), (1, 100.0), (1, 300.0),
          (2, 1e-6), (2, 1e-4), (2, 0.03), (2, 0.3), (2, 0.6)]


def small_state(pkg, scale, ns=3, theta=0.5, seed=7):
    st = pkg.workloads.random_state(12, 40, 30, ns, seed=seed, latlon=True, imp_fac=theta)
    st.dt *= scale
    return st


@pytest.mark.parametrize("dirn,scale", SCALES)
def test_sweeps_match_oracle(pkg, oracle, tr, dirn, scale):
    st = small_state(pkg, scale)
    rho = line_rho(st, dirn)
    want = expected_sweeps(rho)
    name = "diffCN_" + "xyz"[dirn]
    seen = []
    a, b = run(pkg, oracle, tr, st, [("Set_BC", 2), (name,), (name,), (name,)],
               after=lambda T, k: seen.append(T.tune_get("jacobi_last_" + "xyz"[dirn])))
    print(f"rho {rho:.3e}: sweeps {seen[1:]} (want {want})", flush=True)
    assert want > 0, rho
    # the first sweep after the factorisation: dgtsv order
    assert seen[1] == 0 and seen[2] == want and seen[3] == want
    assert_same(a, b, tol=TOL)


@pytest.mark.parametrize("theta", [0.0, 0.5, 1.0])
@pytest.mark.parametrize("nb", [1, 2, 4])
def test_theta_and_bins_per_thread(pkg, oracle, tr, theta, nb):
    st = small_state(pkg, 0.1, ns=5, theta=theta, seed=11)
    seen = []
    calls = [("diffCN_y",), ("diffCN_z",)] * 2
    a, b = run(pkg, oracle, tr, st, calls, nb_diffuse_y=nb, nb_diffuse_z=nb,
               after=lambda T, k: seen.append((T.tune_get("jacobi_last_y"), T.tune_get("jacobi_last_z"))))
    want = (expected_sweeps(line_rho(st, 1)), expected_sweeps(line_rho(st, 2)))
    assert seen[2][0] == want[0] and seen[3][1] == want[1], (seen, want)
    if theta == 0.0:
        assert want == (1, 1)
        assert_same(a, b)  # x = B: bitwise
    else:
        assert_same(a, b, tol=TOL)


def test_subgrid_bounds_and_aloft_mask(pkg, oracle, tr):
    """Active ranges inside the lines (rho over their rows only), lines
    outside the bounds and bins off the IsAloft mask (zero filled)."""
    st = small_state(pkg, 0.1, ns=4, seed=13)
    st.is_aloft = np.array([1, 0, 1, 1], dtype=np.int32)
    st.bounds = (2, 10, 5, 33, 3, 26)
    seen = []
    calls = [("diffCN_y",), ("diffCN_z",)] * 2
    a, b = run(pkg, oracle, tr, st, calls,
               after=lambda T, k: seen.append((T.tune_get("jacobi_last_y"), T.tune_get("jacobi_last_z"))))
    want = (expected_sweeps(line_rho(st, 1, 5, 33)), expected_sweeps(line_rho(st, 2, 3, 26)))
    assert seen[2][0] == want[0] > 0 and seen[3][1] == want[1] > 0, (seen, want)
    assert_same(a, b, tol=TOL)


@pytest.mark.parametrize("case", ["one_fewer", "off", "rho_too_large", "exact_solver"])
def test_dgtsv_order_when_sweeps_do_not_apply(pkg, oracle, tr, case):
    scale = 2000.0 if case == "rho_too_large" else 1e-3
    st = small_state(pkg, scale, seed=17)
    want = expected_sweeps(line_rho(st, 1))
    tune = {}
    if case == "one_fewer":
        assert want > 1
        tune["jacobi_sweeps"] = want - 1
    elif case == "off":
        tune["jacobi_sweeps"] = 0
    elif case == "rho_too_large":
        assert want == 0
    else:
        tune["line_solver"] = "exact"
    seen = []
    a, b = st.copy(), st.copy()
    O = oracle.Oracle(a)
    T = tr.Transport(b)
    T.tune(**({"line_solver": "partitioned", "part_max_line_strided": 3} | tune))
    for _ in range(3):
        O.diffCN_y()
        T.diffCN_y()
        T.synchronize()
        seen.append(T.tune_get("jacobi_last_y"))
    T.sync_to_host()
    T.close()
    assert seen == [0, 0, 0]
    assert_same(a, b)


def test_refactorisation_after_dt_change(pkg, oracle, tr):
    """A new dt refactors the lines: its first sweep runs in dgtsv order,
    the next ones take the sweeps the new rho asks for."""
    st = small_state(pkg, 1e-5, seed=19)
    rho1 = line_rho(st, 1)
    seen = []

    def after(T, k):
        seen.append(T.tune_get("jacobi_last_y"))
        if k == 1:
            T.st.dt *= 100.0

    a, b = st.copy(), st.copy()
    O = oracle.Oracle(a)
    T = tr.Transport(b)
    T.tune(line_solver="partitioned", part_max_line_strided=3)
    for k in range(4):
        O.diffCN_y()
        T.diffCN_y()
        T.synchronize()
        after(T, k)
        if k == 1:
            a.dt = b.dt
    st2 = st.copy()
    st2.dt = b.dt
    rho2 = line_rho(st2, 1)
    T.sync_to_host()
    T.close()
    assert rho2 > rho1
    assert seen == [0, expected_sweeps(rho1), 0, expected_sweeps(rho2)], (seen, rho1, rho2)
    assert_same(a, b, tol=TOL)


@pytest.mark.parametrize("periodic", [False, True])
def test_multistep_with_glue_and_source(pkg, oracle, tr, periodic):
    """Full steps (advection, Set_BC(2), DiffuseHorz, DiffuseVert) with the
    step glue and a vent-column source between them, y and z lines in
    Jacobi sweeps from the second step on."""
    st = pkg.workloads.random_state(16, 40, 24, 3, seed=23, latlon=True, periodic=periodic)
    st.dt *= 0.1
    rng = np.random.default_rng(5)
    a, b = st.copy(), st.copy()
    O = oracle.Oracle(a)
    T = tr.Transport(b)
    T.tune(line_solver="partitioned", part_max_line_strided=3)
    seen = []
    for it in range(1, 5):
        if it > 1:
            O.step_glue()
            T.step_glue()
            flux = np.asfortranarray(rng.random((st.nz + 1, st.ns)) * 20.0)
            O.source_column_add(5, 17, flux)
            T.source_column_add(5, 17, flux)
        O.transport_step(it, True)
        T.transport_step(it, True)
        T.synchronize()
        seen.append((T.tune_get("jacobi_last_y"), T.tune_get("jacobi_last_z")))
    T.sync_to_host()
    T.close()
    print("sweeps per step (y, z):", seen, flush=True)
    assert all(y > 0 and z > 0 for y, z in seen[1:]), seen
    assert_same(a, b, names=ARRAYS, tol=TOL)


def x_state(pkg, nx, ny, nz, ns, scale, seed, periodic=False, theta=0.5):
    """Lon-lat grid from about 30-40 deg N: the rows' rho grows with
    latitude (1 / cos^2), so a dt scale can put the line split anywhere."""
    st = pkg.workloads.random_state(nx, ny, nz, ns, seed=seed, latlon=True, periodic=periodic, imp_fac=theta)
    st.dt *= scale
    return st


# (periodic, dt scale): all lines (6 and 8 sweeps), 129 of 240 (the
# others need more than 12), none; periodic: all, 92 of 240, none
X_SPLITS = [(False, 30.0), (False, 150.0), (False, 2000.0), (False, 1e4), (True, 1e4), (True, 2e6),
            (True, 3e6)]


@pytest.mark.parametrize("periodic,scale", X_SPLITS)
def test_x_lines_split(pkg, oracle, tr, periodic, scale):
    """x lines (jacobi_x=2: any length): every line, some, or none in
    Jacobi sweeps, the others in the partitioned solves."""
    st = x_state(pkg, 24, 40, 6, 3, scale, seed=29, periodic=periodic)
    nj, nl, want = x_split(st)
    seen = []
    a, b = run(pkg, oracle, tr, st, [("Set_BC", 2), ("diffCN_x",), ("diffCN_x",), ("diffCN_x",)], jacobi_x=2,
               after=lambda T, k: seen.append(T.tune_get("jacobi_last_x")))
    print(f"{nj} of {nl} lines in {want} sweeps; seen {seen}", flush=True)
    assert seen[1] == 0 and seen[2] == seen[3] == want, (seen, want)
    assert_same(a, b, tol=TOL)


def test_x_lines_split_covers_both_kinds(pkg):
    """The scales above put some lines on each side (checked on the CPU)."""
    for periodic in (False, True):
        kinds = set()
        for p, scale in X_SPLITS:
            if p == periodic:
                nj, nl, _ = x_split(x_state(pkg, 24, 40, 6, 3, scale, seed=29, periodic=p))
                kinds.add("all" if nj == nl else "none" if nj == 0 else "some")
        assert kinds == {"all", "some", "none"}, (periodic, kinds)


@pytest.mark.parametrize("theta,scale", [(0.5, 2000.0), (1.0, 1000.0)])
@pytest.mark.parametrize("nb", [1, 2, 4])
@pytest.mark.parametrize("nx", [300, 600])
def test_long_x_lines(pkg, oracle, tr, theta, scale, nb, nx):
    """Lines of 300 and 600 cells (default jacobi_x=1: lines of 256 cells
    and more; one and two waves of the Jacobi kernel, one and three parts
    of the partitioned solves), 51 or 59 of 72 lines in Jacobi sweeps, subgrid
    bounds and an IsAloft mask, 1, 2 and 4 bins per thread."""
    st = x_state(pkg, nx, 24, 3, 5, scale, seed=31, theta=theta)
    st.is_aloft = np.array([1, 1, 0, 1, 1], dtype=np.int32)
    st.bounds = (3, nx - 10, 2, 23, 1, 3)
    nj, nl, want = x_split(st, 3, nx - 10)
    seen = []
    a, b = run(pkg, oracle, tr, st, [("diffCN_x",)] * 3, nb_diffuse_x=nb,
               after=lambda T, k: seen.append(T.tune_get("jacobi_last_x")))
    print(f"{nj} of {nl} lines in {want} sweeps; seen {seen}", flush=True)
    assert 0 < nj < nl
    assert seen == [0, want, want], (seen, want)
    assert_same(a, b, tol=TOL)


def test_x_lines_sweep_counts_per_line(pkg, oracle, tr):
    """Rows from 30 to 80 deg: lines needing 6 to 12 sweeps and lines
    needing more in one sweep (each takes its own number); with the cap at
    8 the lines needing 9 to 12 go to the partitioned solves too."""
    st = x_state(pkg, 24, 100, 4, 2, 100.0, seed=29)
    sw = np.vectorize(lambda r: expected_sweeps(r, MAX_SWEEPS_X))(per_line_rho(st, 0))
    counts = {int(v): int((sw == v).sum()) for v in np.unique(sw)}
    print("lines per sweep count:", counts, flush=True)
    assert set(counts) == {0, 6, 7, 8, 9, 10, 11, 12}, counts
    for cap in (12, 8):
        seen = []
        a, b = run(pkg, oracle, tr, st, [("diffCN_x",)] * 3, jacobi_x=2, jacobi_sweeps=cap,
                   after=lambda T, k: seen.append(T.tune_get("jacobi_last_x")))
        want = min(max(counts), cap)
        assert seen == [0, want, want], (cap, seen, want)
        assert_same(a, b, tol=TOL)


def test_short_x_lines_keep_the_partitioned_solves(pkg, oracle, tr):
    """jacobi_x=1 (default): x lines shorter than 256 cells stay in the
    partitioned solves."""
    st = x_state(pkg, 24, 40, 6, 3, 30.0, seed=29)
    seen = []
    a, b = run(pkg, oracle, tr, st, [("diffCN_x",)] * 3,
               after=lambda T, k: seen.append(T.tune_get("jacobi_last_x")))
    assert seen == [0, 0, 0]
    assert_same(a, b, tol=TOL)


# ---- the adjoint: the transpose of the Jacobi map (same sweeps on A^T =
# K A K^-1: lambda <- K B p(A) K^-1 lambda), pinned by <F q, w> = <q, F^T w>
# against the oracle's forward map (whose exact solves the sweeps match to
# 2^-52 relative)

def adjoint_jacobi(tr, st, lam, dirn, **tune):
    """F^T lam through the library's adjoint after one warm-up call (which
    factors the lines; the factors' ratios then reach the host), and the
    sweeps the second call reports."""
    T = tr.Transport(st.copy())
    T.tune(line_solver="partitioned", **tune)
    T.adj_init(0)
    T.adj_set(lam)
    T.adj_diffuse(dirn, True)
    T.synchronize()
    T.adj_set(lam)
    T.adj_diffuse(dirn, True)
    g = T.adj_get()
    sweeps = T.tune_get("jacobi_last_" + "xyz"[dirn])
    T.close()
    return g, sweeps


@pytest.mark.parametrize("dirn,scale", [(1, 1e-3), (1, 1.0), (1, 300.0), (2, 1e-4), (2, 0.3), (2, 0.6)])
def test_adjoint_sweeps(pkg, oracle, tr, dirn, scale):
    from test_gpu_adjoint import check_dot, forward, rand_lambda
    st = small_state(pkg, scale)
    want = expected_sweeps(line_rho(st, dirn))
    Fq = forward(oracle, st, [("diffCN_" + "xyz"[dirn],)])
    lam = rand_lambda(st, 70 + dirn)
    g, sweeps = adjoint_jacobi(tr, st, lam, dirn, part_max_line_strided=3)
    assert sweeps == want > 0, (sweeps, want)
    check_dot(st, Fq, lam, g)


def test_adjoint_subgrid_mask_and_theta(pkg, oracle, tr):
    from test_gpu_adjoint import check_dot, forward, rand_lambda
    for theta in (0.0, 1.0):
        st = small_state(pkg, 0.1, ns=4, theta=theta, seed=13)
        st.is_aloft = np.array([1, 0, 1, 1], dtype=np.int32)
        st.bounds = (2, 10, 5, 33, 3, 26)
        for dirn in (1, 2):
            Fq = forward(oracle, st, [("diffCN_" + "xyz"[dirn],)])
            lam = rand_lambda(st, 80 + dirn)
            g, sweeps = adjoint_jacobi(tr, st, lam, dirn, part_max_line_strided=3)
            assert sweeps > 0
            check_dot(st, Fq, lam, g)


@pytest.mark.parametrize("periodic,scale", [(False, 30.0), (False, 2000.0), (True, 2e6)])
def test_adjoint_x_lines_split(pkg, oracle, tr, periodic, scale):
    """x lines of one wave (jacobi_x=2): the Jacobi lines' adjoint, the
    others in the partitioned adjoint (which skips the Jacobi lines)."""
    from test_gpu_adjoint import check_dot, forward, rand_lambda
    st = x_state(pkg, 24, 40, 6, 3, scale, seed=29, periodic=periodic)
    nj, nl, want = x_split(st)
    Fq = forward(oracle, st, [("diffCN_x",)])
    lam = rand_lambda(st, 90)
    g, sweeps = adjoint_jacobi(tr, st, lam, 0, jacobi_x=2)
    assert sweeps == want > 0, (sweeps, want)
    check_dot(st, Fq, lam, g)


def test_adjoint_long_x_lines(pkg, oracle, tr):
    """600-cell x lines (default jacobi_x=1): 51 of 72 lines in the Jacobi
    adjoint, the others in the multi-wave partitioned adjoint (which skips
    the Jacobi lines); subgrid and a mask."""
    from test_gpu_adjoint import check_dot, forward, rand_lambda
    st = x_state(pkg, 600, 24, 3, 5, 2000.0, seed=31)
    st.is_aloft = np.array([1, 1, 0, 1, 1], dtype=np.int32)
    st.bounds = (3, 590, 2, 23, 1, 3)
    nj, nl, want = x_split(st, 3, 590)
    assert 0 < nj < nl
    Fq = forward(oracle, st, [("diffCN_x",)])
    lam = rand_lambda(st, 91)
    g, sweeps = adjoint_jacobi(tr, st, lam, 0)
    assert sweeps == want, (sweeps, want)
    check_dot(st, Fq, lam, g)
