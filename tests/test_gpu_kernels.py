"""The reference's KernelContractSuite (KernelsSuite.scala:19-186) held
against the HIP backend in both precisions, plus random SpMV/axpby/dot
against numpy. Tolerance follows the declared precision: 1e-12 for fp64,
1e-6 otherwise (KernelsSuite.scala:53-56)."""
import numpy as np
import pytest

from noaidi_amd import prima

pytestmark = pytest.mark.gpu


@pytest.fixture(params=[64, 32], ids=["f64", "f32"])
def k(request, hip64, hip32):
    return hip64 if request.param == 64 else hip32


def tol(k):
    return 1e-12 if k.capabilities["supportsFloat64"] else 1e-6


def close(k, a, e):
    a = np.asarray(a)
    e = np.asarray(e, dtype=float)
    assert a.shape == e.shape
    assert np.all(np.abs(a - e) <= tol(k) * np.maximum(1.0, np.abs(e))), (a, e)


def test_round_trip_exact(k):
    data = np.array([1.0, -2.5, 3.75, 0.0])
    assert list(k.download(k.upload(data))) == list(data)


def test_upload_copies(k):
    data = np.array([1.0, 2.0])
    v = k.upload(data)
    data[0] = 99.0
    assert k.download(v)[0] == 1.0


def test_allocate_zero(k):
    assert list(k.download(k.allocate(5))) == [0.0] * 5


def test_spmv_dense_product(k):
    m = k.uploadMatrix(2, 3, [0, 2, 4], [0, 1, 1, 2], [1.0, 2.0, -1.0, 3.0])
    x = k.upload([2.0, 3.0, 4.0])
    out = k.allocate(2)
    k.spmv(m, x, out)
    close(k, k.download(out), [8.0, 9.0])


def test_axpby_aliased(k):
    x = k.upload([1.0, 2.0, 3.0])
    y = k.upload([10.0, 20.0, 30.0])
    out = k.allocate(3)
    k.axpby(2.0, x, -1.0, y, out)
    close(k, k.download(out), [-8.0, -16.0, -24.0])
    k.axpby(2.0, x, -1.0, y, y)
    close(k, k.download(y), [-8.0, -16.0, -24.0])


def test_dot_and_norm(k):
    x = k.upload([1.0, 2.0, 3.0])
    y = k.upload([4.0, -5.0, 6.0])
    assert abs(k.dot(x, y) - 12.0) <= tol(k)
    assert abs(k.squaredNorm(x) - 14.0) <= tol(k)
    assert k.squaredNorm(k.allocate(3)) == 0.0


def test_primal_step_box(k):
    z = k.upload([0.0, 0.0, 0.0])
    out = k.allocate(3)
    k.primalStep(z, k.upload([0.0] * 3), k.upload([1.0, -1.0, 0.0]), k.upload([-0.5, 0.0, -1.0]),
                 k.upload([1.0, 0.5, 1.0]), 1.0, out)
    close(k, k.download(out), [-0.5, 0.5, 0.0])


def test_primal_step_infinite_bounds(k):
    out = k.allocate(2)
    k.primalStep(k.upload([0.0, 0.0]), k.upload([0.0, 0.0]), k.upload([1.0, -1.0]),
                 k.upload([-np.inf, 0.0]), k.upload([0.0, np.inf]), 2.0, out)
    close(k, k.download(out), [-2.0, 2.0])


def test_dual_step_cone(k):
    out = k.allocate(4)
    k.dualStep(k.upload([0.0] * 4), k.upload([0.0] * 4), k.upload([-5.0, 5.0, -5.0, 5.0]), 1.0,
               2, out)
    close(k, k.download(out), [-5.0, 5.0, 0.0, 5.0])


def test_capabilities(k):
    c = k.capabilities
    assert c["name"]
    assert c["requiresDoublePrecisionRefinement"] == (not c["supportsFloat64"])
    assert c["supportsFloat64"] == (k.precision == 64)


def test_float32_narrowing(hip32):
    # Float32KernelsSuite (KernelsSuite.scala:208-245)
    v = hip32.download(hip32.upload([1.0 + 2.0 ** -30, np.inf, -np.inf, 1e300]))
    assert v[0] == 1.0 and v[1] == np.inf and v[2] == -np.inf and v[3] == np.inf


@pytest.mark.parametrize("rows,cols,dens", [(1, 1, 1.0), (37, 53, 0.2), (500, 300, 0.02),
                                            (64, 4000, 0.5)])
def test_random_spmv_and_vector_ops(k, rows, cols, dens):
    import scipy.sparse as sp
    rng = np.random.default_rng(rows * 7 + cols)
    A = sp.random(rows, cols, density=dens, random_state=rng, format="csr")
    A.sort_indices()
    x = rng.standard_normal(cols)
    m = k.uploadMatrix(rows, cols, A.indptr, A.indices, A.data)
    xv, out = k.upload(x), k.allocate(rows)
    k.spmv(m, xv, out)
    ref = A @ x
    scale = np.abs(A) @ np.abs(x) + 1.0
    assert np.all(np.abs(k.download(out) - ref) <= 10 * tol(k) * scale)
    y = rng.standard_normal(cols)
    yv = k.upload(y)
    assert abs(k.dot(xv, yv) - x @ y) <= 10 * tol(k) * (np.abs(x) @ np.abs(y) + 1)
    o2 = k.allocate(cols)
    k.scale(0.5, xv, o2)
    close(k, k.download(o2), 0.5 * (x.astype(np.float32) if k.precision == 32 else x))


def _wide_matrix(rows, cols, seed):
    # lognormal row lengths (mean about 20, a few rows longer than 512 so the
    # panel schedule's separate long-row pass runs too), uniform columns
    import scipy.sparse as sp
    rng = np.random.default_rng(seed)
    lens = np.minimum(np.exp(rng.normal(np.log(20.0) - 0.5, 1.0, rows)).astype(np.int64) + 1, 2048)
    lens[:3] = [600, 1500, 2048]
    ptr = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    col = rng.integers(0, cols, int(ptr[-1])).astype(np.int32)
    val = rng.standard_normal(int(ptr[-1]))
    A = sp.csr_matrix((val, col, ptr), shape=(rows, cols))
    return A, ptr, col, val, rng


@pytest.mark.parametrize("prec", [32, 64])
def test_panel_schedule_matches_host(hip32, hip64, prec):
    # x larger than an XCD's L2 (2^20 columns in f32, 2^19 in f64): the panel
    # layout is built at upload; every schedule agrees with the f64 host product
    k = hip32 if prec == 32 else hip64
    cols = (1 << 20) + 12345 if prec == 32 else (1 << 19) + 777
    A, ptr, col, val, rng = _wide_matrix(40_000, cols, prec)
    x = rng.standard_normal(cols)
    m = k.uploadMatrix(A.shape[0], cols, ptr, col, val)
    xv, out = k.upload(x), k.allocate(A.shape[0])
    ref = A @ x
    scale = np.abs(A) @ np.abs(x) + 1.0
    t = 1e-12 if prec == 64 else 1e-6
    results = {}
    for sch in range(1, 11):
        k.set_schedule(m, sch)
        k.spmv(m, xv, out)
        got = k.download(out)
        assert np.all(np.abs(got - ref) <= 10 * t * scale), sch
        results[sch] = got
    # deterministic: the panel schedule gives the same bits twice
    k.set_schedule(m, 9)
    k.spmv(m, xv, out)
    assert np.array_equal(k.download(out), results[9])
    choice, ms = k.autotune(m, 3)
    assert 1 <= choice <= 10 and ms > 0


@pytest.mark.parametrize("prec", [32, 64])
def test_row_block_schedules_cover_long_and_short_rows(hip32, hip64, prec):
    # short dispatch-like rows, a run of empty rows, rows exactly at the small
    # schedule's block capacity (2,048) and rows beyond it (summed by the
    # whole workgroup), more rows than one small block holds: the row-block
    # schedules (solver-sized and small) agree with the f64 host product and
    # give the same bits twice
    import scipy.sparse as sp
    k = hip32 if prec == 32 else hip64
    rng = np.random.default_rng(prec)
    rows, cols = 7000, 50_000
    lens = rng.integers(3, 7, rows)
    lens[100:140] = 0
    lens[[7, 2000]] = 2048
    lens[[11, 3001, 6999]] = [2049, 5000, 9000]
    ptr = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    col = rng.integers(0, cols, int(ptr[-1])).astype(np.int32)
    val = rng.standard_normal(int(ptr[-1]))
    A = sp.csr_matrix((val, col, ptr), shape=(rows, cols))
    x = rng.standard_normal(cols)
    m = k.uploadMatrix(rows, cols, ptr, col, val)
    xv, out = k.upload(x), k.allocate(rows)
    ref = A @ x
    scale = np.abs(A) @ np.abs(x) + 1.0
    t = 1e-12 if prec == 64 else 1e-6
    for sch in (8, 10):
        k.set_schedule(m, sch)
        k.spmv(m, xv, out)
        got = k.download(out)
        assert np.all(np.abs(got - ref) <= 10 * t * scale), sch
        k.spmv(m, xv, out)
        assert np.array_equal(k.download(out), got), sch


def test_panel_schedule_needs_a_wide_matrix(hip64):
    m = hip64.uploadMatrix(2, 3, [0, 1, 2], [0, 2], [1.0, 2.0])
    with pytest.raises(ValueError):
        hip64.set_schedule(m, 9)
    with pytest.raises(ValueError):
        hip64.set_schedule(m, 11)
