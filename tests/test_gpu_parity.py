"""GPU parity: the HIP kernels through the C ABI against the CPU oracle on the
same seeded inputs.  Integer, Bool and min-plus results must be bit-exact;
fp32 plus-times must be within 1e-5 relative of an FP64 product (and of the
oracle's fp32 result)."""
import numpy as np
import pytest
import scipy.sparse as sp

from conftest import print_case, random_csr
import graphs as G

pytestmark = pytest.mark.gpu

INT32_MAX = np.iinfo(np.int32).max
RTOL = 1e-5  # fp32 plus-times vs FP64 product, relative


@pytest.fixture(scope="module")
def env(oracle):
    import torch
    import gpugraphs as gg
    assert torch.cuda.is_available()
    gg.lib()
    return gg, torch, oracle


def dev(torch, a):
    a = np.ascontiguousarray(a)
    return torch.from_numpy(a).cuda()


def host(t):
    return t.cpu().numpy()


# sell_relabel_sorted: each row's entries by increasing internal column
# sell_relabel_rcm: equal-length rows in reverse Cuthill-McKee order
FORMATS = ["csr", "sell_ref", "sell_hybrid", "sell_relabel", "sell_relabel_sorted",
           "sell_relabel_rcm"]


def make(gg, S, fmt):
    R = gg.ROCBackend(0)
    if fmt == "csr":
        return gg.SparseGPUMatrixCSR(S, R)
    if fmt == "sell_ref":
        return gg.SparseGPUMatrixSELL(S, R, 32)
    if fmt.startswith("sell_relabel"):
        # degree-relabelled rows (internal column ids) with a heavy partition
        if S.shape[0] != S.shape[1]:
            pytest.skip("relabelling needs a square matrix")
        return gg.SparseGPUMatrixSELL(S, R, 64, heavy_threshold=8, heavy_seg_len=16, relabel=True,
                                      sort_entries=fmt.endswith(("_sorted", "_rcm")),
                                      rcm=fmt.endswith("_rcm"))
    # sigma-sorted C=64 slices with a heavy partition cut into 16-entry segments
    return gg.SparseGPUMatrixSELL(S, R, 64, sigma=64, heavy_threshold=8, heavy_seg_len=16)


def orc_mat(oracle, S, fmt):
    A = oracle.from_scipy_csr(S)
    return oracle.sell_from_csr(A, 32) if fmt != "csr" else A


@pytest.mark.parametrize("fmt", FORMATS)
@pytest.mark.parametrize("m,n,p", [(10, 10, 0.5), (1000, 1000, 0.2), (300, 700, 0.05),
                                   (4096, 4096, 0.004)])
def test_spmv_plus_times_f32(env, fmt, m, n, p):
    gg, torch, orc = env
    print_case("spmv plus-times", fmt, m, n, p)
    rng = np.random.default_rng(m + n)
    S = random_csr(rng, m, n, p)
    x = rng.random(n, dtype=np.float32)
    A = make(gg, S, fmt)
    c = dev(torch, np.zeros(m, np.float32))
    gg.gpu_spmv_(c, A, dev(torch, x))
    torch.cuda.synchronize()
    got = host(c)
    ref64 = S.astype(np.float64) @ x.astype(np.float64)
    np.testing.assert_allclose(got, ref64, rtol=RTOL, atol=1e-7)
    co = np.zeros(m, np.float32)
    orc.spmv(co, orc_mat(orc, S, "csr"), x)
    np.testing.assert_allclose(got, co, rtol=RTOL, atol=1e-7)


@pytest.mark.parametrize("fmt", FORMATS)
def test_spmv_power_law_rows_f32(env, fmt):
    """Long rows (hubs) accumulate in segments; the tolerance still holds."""
    gg, torch, orc = env
    from gpugraphs import generators as GEN
    print_case("spmv power law", fmt)
    R = GEN.rmat(15, seed=3, dtype=np.float32, values=GEN.UNIFORM01)
    S = R.to_scipy()
    x = np.random.default_rng(1).random(R.n, dtype=np.float32)
    if fmt == "sell_hybrid":
        A = R.to_sell(gg.ROCBackend(0), 64, 4096, 256, 512)
    elif fmt.startswith("sell_relabel"):
        A = R.to_sell(gg.ROCBackend(0), 64, 1, 256, 512, relabel=True,
                      sort_entries=fmt.endswith("_sorted"))
    elif fmt == "sell_ref":
        A = R.to_sell(gg.ROCBackend(0), 64, 1, 0)
    else:
        A = R.to_csr(gg.ROCBackend(0))
    c = dev(torch, np.zeros(R.m, np.float32))
    gg.gpu_spmv_(c, A, dev(torch, x))
    torch.cuda.synchronize()
    np.testing.assert_allclose(host(c), S.astype(np.float64) @ x.astype(np.float64), rtol=RTOL,
                               atol=1e-7)


SEMIRINGS = [
    # (add, mul, accum, A dtype, x dtype, out dtype)
    ("bor", "band", "second", np.bool_, np.bool_, np.bool_),
    ("bor", "band", "second", np.int32, np.int32, np.int32),
    ("min", "safe_add", "min", np.int32, np.int32, np.int32),
    ("min", "safe_add", "min", np.float32, np.float32, np.float32),
    ("min", "safe_add", "min", np.bool_, np.int32, np.int32),
    ("min", "add", "second", np.int32, np.int32, np.int32),
    ("add", "mul", "second", np.int32, np.int32, np.int32),
    ("add", "mul", "add", np.int64, np.int64, np.int64),
    ("any", "secondi", "second", np.int32, np.int32, np.int32),
    ("lor", "land", "second", np.bool_, np.bool_, np.bool_),
]


def _ops(gg, name):
    return {"bor": gg.GPUGraphs_bor, "band": gg.GPUGraphs_band, "min": gg.GPUGraphs_min,
            "safe_add": gg.GPUGraphs_safe_add, "add": gg.GPUGraphs_add, "mul": gg.GPUGraphs_mul,
            "any": gg.GPUGraphs_any, "secondi": gg.GPUGraphs_secondi,
            "second": gg.GPUGraphs_second, "lor": gg.GPUGraphs_lor,
            "land": gg.GPUGraphs_land}[name]


def _values(rng, dt, size, lo=1, hi=1000):
    if dt == np.bool_:
        return rng.random(size) < 0.5
    if dt == np.float32:
        return rng.integers(lo, hi, size).astype(np.float32)
    return rng.integers(lo, hi, size).astype(dt)


@pytest.mark.parametrize("fmt", FORMATS)
@pytest.mark.parametrize("sr", SEMIRINGS, ids=lambda s: "-".join(str(x) for x in s[:3]) + f"-{np.dtype(s[3]).name}")
@pytest.mark.parametrize("variant", ["plain", "mask", "range", "range_mask"])
def test_spmv_semirings_exact(env, fmt, sr, variant):
    gg, torch, orc = env
    add, mul, acc, ta, tx, tc = sr
    print_case("spmv exact", fmt, sr, variant)
    rng = np.random.default_rng(hash((fmt, variant, sr[:3])) % 2**32)
    m = n = 700
    S = random_csr(rng, m, n, 0.02)
    S.data = _values(rng, ta, S.nnz)
    if ta == np.bool_:
        S.data[:] = True  # sparse() keeps only nonzeros
    S = S.astype(ta)
    x = _values(rng, tx, n)
    if add == "any":
        x = (rng.random(n) < 0.3).astype(tx)
    c0 = _values(rng, tc, m) if acc != "second" or variant != "plain" else np.zeros(m, tc)
    mask = (rng.random(m) < 0.5).astype(np.int32) if "mask" in variant else None
    rng_ = range(101, 401) if "range" in variant else None
    A = make(gg, S, fmt)
    c = dev(torch, c0.copy())
    kw = dict(mul=_ops(gg, mul), add=_ops(gg, add), accum=_ops(gg, acc))
    if mask is not None:
        kw["mask"] = dev(torch, mask)
    if rng_ is not None:
        kw["range_"] = rng_
    gg.gpu_spmv_(c, A, dev(torch, x), **kw)
    torch.cuda.synchronize()
    co = c0.copy()
    if co.dtype == np.bool_:
        co = co.view(np.uint8)
    orc.spmv(co, orc_mat(orc, S, fmt), x, add=getattr(orc, "OP_" + add.upper()),
             mul=getattr(orc, "OP_" + mul.upper()), accum=getattr(orc, "OP_" + acc.upper()),
             mask=mask, range_=None if rng_ is None else (rng_.start, len(rng_)))
    got = host(c)
    if got.dtype == np.bool_:
        got = got.view(np.uint8)
    np.testing.assert_array_equal(got, co)


@pytest.mark.parametrize("sr", SEMIRINGS, ids=lambda s: "-".join(str(x) for x in s[:3]) + f"-{np.dtype(s[3]).name}")
def test_spmv_sparse_mask_rows(env, sr):
    """A SparseGPUVector mask lists the rows to compute (the CSR kernel of
    src/spmv/CSR.jl:90-115): listed rows equal the unmasked product's rows
    (the generic fold, so any = max as in that kernel), the others keep C."""
    gg, torch, orc = env
    add, mul, acc, ta, tx, tc = sr
    print_case("spmv sparse mask", sr)
    rng = np.random.default_rng(hash(sr[:3]) % 2**32)
    m = n = 700
    S = random_csr(rng, m, n, 0.02)
    S.data = _values(rng, ta, S.nnz)
    if ta == np.bool_:
        S.data[:] = True
    S = S.astype(ta)
    x = _values(rng, tx, n)
    if add == "any":
        x = (rng.random(n) < 0.3).astype(tx)
    c0 = _values(rng, tc, m)
    rows = np.sort(rng.choice(m, 250, replace=False)).astype(np.int32) + 1
    rows = rng.permutation(rows)
    kw = dict(mul=_ops(gg, mul), add=_ops(gg, add), accum=_ops(gg, acc))
    A = make(gg, S, "csr")
    mk = gg.SparseGPUVector(m, rows, np.ones(len(rows), np.int32), gg.ROCBackend(0))
    c = dev(torch, c0.copy())
    gg.gpu_spmv_(c, A, dev(torch, x), mask=mk, **kw)
    torch.cuda.synchronize()
    full = c0.copy()
    if full.dtype == np.bool_:
        full = full.view(np.uint8)
    orc.spmv(full, orc_mat(orc, S, "csr"), x, add=getattr(orc, "OP_" + add.upper()),
             mul=getattr(orc, "OP_" + mul.upper()), accum=getattr(orc, "OP_" + acc.upper()))
    want = c0.copy()
    if want.dtype == np.bool_:
        want = want.view(np.uint8)
    want[rows - 1] = full[rows - 1]
    got = host(c)
    if got.dtype == np.bool_:
        got = got.view(np.uint8)
    np.testing.assert_array_equal(got, want)
    # an empty list leaves C alone; SELL has no row-list kernel; lengths are checked
    empty = gg.SparseGPUVector(m, np.zeros(0, np.int32), np.zeros(0, np.int32), gg.ROCBackend(0))
    gg.gpu_spmv_(c, A, dev(torch, x), mask=empty, **kw)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(got, host(c).view(np.uint8) if host(c).dtype == np.bool_ else host(c))
    with pytest.raises(gg.MethodError):
        gg.gpu_spmv_(c, make(gg, S, "sell_ref"), dev(torch, x), mask=mk, **kw)
    with pytest.raises(gg.DimensionMismatch):
        gg.gpu_spmv_(c, A, dev(torch, x), mask=gg.SparseGPUVector(m + 1, rows, np.ones(len(rows), np.int32),
                                                                    gg.ROCBackend(0)), **kw)


def test_spmv_errors(env):
    gg, torch, orc = env
    A = gg.SparseGPUMatrixCSR(sp.random(30, 40, density=0.1, format="csr", dtype=np.float32),
                              gg.ROCBackend(0))
    f32 = lambda k: dev(torch, np.zeros(k, np.float32))
    with pytest.raises(gg.DimensionMismatch):
        gg.gpu_spmv_(f32(30), A, f32(41))
    with pytest.raises(gg.DimensionMismatch):
        gg.gpu_spmv_(f32(31), A, f32(40))
    with pytest.raises(gg.ArgumentError):
        gg.gpu_spmv_(dev(torch, np.zeros(30, np.int32)), A, f32(40))
    with pytest.raises(gg.DimensionMismatch):
        gg.gpu_spmv_(f32(30), A, f32(40), mask=dev(torch, np.ones(29, np.int32)))
    with pytest.raises(gg.BoundsError):
        gg.gpu_spmv_(f32(30), A, f32(40), range_=range(20, 32))
    with pytest.raises(gg.MethodError):
        gg.gpu_spmv_(f32(30), A, f32(40), add=gg.GPUGraphs_max, mul=gg.GPUGraphs_min)
    with pytest.raises(gg.ArgumentError):
        gg.gpu_spmv_(f32(30), A, np.zeros(40, np.float32))  # host vector: not the same backend


def _cm(torch, a):
    """Column-major (Julia layout) device copy of a 2-D array."""
    return dev(torch, np.ascontiguousarray(a.T)).T


@pytest.mark.parametrize("fmt", FORMATS)
@pytest.mark.parametrize("k", [1, 5, 20, 32, 33, 64, 96])
@pytest.mark.parametrize("layout", ["col", "row"])
def test_spmm_plus_times_and_min_plus(env, fmt, k, layout):
    gg, torch, orc = env
    print_case("spmm", fmt, k, layout)
    rng = np.random.default_rng(k * 3 + len(fmt))
    S = random_csr(rng, 400, 400, 0.03)
    B = rng.random((400, k), dtype=np.float32)
    A = make(gg, S, fmt)
    mk = (lambda a: _cm(torch, a)) if layout == "col" else (lambda a: dev(torch, a))
    Cd = mk(np.zeros((400, k), np.float32))
    gg.gpu_spmm_(Cd, A, mk(B))
    torch.cuda.synchronize()
    np.testing.assert_allclose(host(Cd), S.astype(np.float64) @ B.astype(np.float64), rtol=RTOL,
                               atol=1e-7)
    # min-plus with integer weights, exact against the oracle
    Si = S.copy()
    Si.data = rng.integers(1, 1000, Si.nnz).astype(np.float32)
    Bi = rng.integers(0, 5000, (400, k)).astype(np.float32)
    Bi[rng.random((400, k)) < 0.3] = np.inf
    Ai = make(gg, Si, fmt)
    C0 = rng.integers(0, 9000, (400, k)).astype(np.float32)
    Cd = mk(C0)
    gg.gpu_spmm_(Cd, Ai, mk(Bi), mul=gg.GPUGraphs_safe_add, add=gg.GPUGraphs_min,
                 accum=gg.GPUGraphs_min)
    torch.cuda.synchronize()
    Co = np.asfortranarray(C0.copy())
    orc.spmm(Co, orc_mat(orc, Si, fmt), np.asfortranarray(Bi), add=orc.OP_MIN,
             mul=orc.OP_SAFE_ADD, accum=orc.OP_MIN)
    np.testing.assert_array_equal(host(Cd), Co)


@pytest.mark.parametrize("k", [6, 8, 32])
@pytest.mark.parametrize("layout", ["col", "row"])
def test_spmm_csr_mask_range_and_asserts(env, k, layout):
    """CSR SpMM with a row mask or a row range (one lane per column for k = 6,
    16-byte column groups for 8 and 32), exact against the oracle."""
    gg, torch, orc = env
    rng = np.random.default_rng(7 + k)
    S = random_csr(rng, 300, 300, 0.05)
    S.data = rng.integers(1, 100, S.nnz).astype(np.float32)
    B = rng.integers(0, 100, (300, k)).astype(np.float32)
    A = gg.SparseGPUMatrixCSR(S, gg.ROCBackend(0))
    mask = (rng.random(300) < 0.4).astype(np.int32)
    mk = (lambda a: _cm(torch, a)) if layout == "col" else (lambda a: dev(torch, a))
    print_case("spmm csr mask range", k, layout)
    for kw, okw in [({"mask": dev(torch, mask)}, {"mask": mask}),
                    ({"range_": range(50, 151)}, {"range_": (50, 101)}), ({}, {})]:
        C0 = np.full((300, k), -1.0, np.float32)
        Cd = mk(C0)
        gg.gpu_spmm_(Cd, A, mk(B), **kw)
        torch.cuda.synchronize()
        Co = np.asfortranarray(C0.copy())
        orc.spmm(Co, orc_mat(orc, S, "csr"), np.asfortranarray(B), **okw)
        np.testing.assert_array_equal(host(Cd), Co)
    B = B[:, :6].copy()
    with pytest.raises(AssertionError):
        gg.gpu_spmm_(_cm(torch, np.zeros((300, 5), np.float32)), A, _cm(torch, B))
    with pytest.raises(AssertionError):
        gg.gpu_spmm_(_cm(torch, np.zeros((300, 6), np.float32)), A, _cm(torch, B),
                     range_=range(250, 320))


def test_csc_spmv(env):
    gg, torch, orc = env
    rng = np.random.default_rng(2)
    for m, p in ((10, 0.5), (1000, 0.2)):  # test/spmv.jl:16-68 sizes
        S = random_csr(rng, m, m, p)
        x = rng.random(m, dtype=np.float32)
        A = gg.SparseGPUMatrixCSC(S, gg.ROCBackend(0))
        c = dev(torch, np.zeros(m, np.float32))
        gg.gpu_spmv_(c, A, dev(torch, x))
        torch.cuda.synchronize()
        np.testing.assert_allclose(host(c), S.astype(np.float64) @ x, rtol=RTOL, atol=1e-6)


@pytest.mark.parametrize("layout", ["col", "row"])
def test_csc_spmm(env, layout):
    """CSC SpMM, the k-column form of the CSC SpMV (atomic adds into C): fp32
    within the tolerance of an FP64 product, Int32 exact; Julia's
    column-major and row-major operands; C accumulates."""
    gg, torch, orc = env
    rng = np.random.default_rng(3)
    conv = (lambda a: _cm(torch, a)) if layout == "col" else (lambda a: dev(torch, a))
    for m, n, k, p in ((10, 12, 3, 0.5), (1000, 900, 20, 0.05)):
        S = random_csr(rng, m, n, p)
        print_case("csc spmm", layout, m, n, k)
        B = rng.random((n, k), dtype=np.float32)
        A = gg.SparseGPUMatrixCSC(S, gg.ROCBackend(0))
        C0 = rng.random((m, k), dtype=np.float32)
        c = conv(C0.copy())
        gg.gpu_spmm_(c, A, conv(B))
        torch.cuda.synchronize()
        np.testing.assert_allclose(host(c), C0 + S.astype(np.float64) @ B.astype(np.float64),
                                   rtol=RTOL, atol=1e-5)
        Si = S.copy()
        Si.data = rng.integers(1, 50, S.nnz).astype(np.float32)
        Si = Si.astype(np.int32)
        Bi = rng.integers(0, 50, (n, k)).astype(np.int32)
        Ai = gg.SparseGPUMatrixCSC(Si, gg.ROCBackend(0))
        ci = conv(np.zeros((m, k), np.int32))
        gg.gpu_spmm_(ci, Ai, conv(Bi))
        torch.cuda.synchronize()
        np.testing.assert_array_equal(host(ci), (Si.astype(np.int64) @ Bi).astype(np.int32))
    with pytest.raises(gg.ArgumentError):
        gg.gpu_spmm_(conv(np.zeros((m, k), np.float32)), Ai, conv(Bi))
    with pytest.raises(AssertionError):
        gg.gpu_spmm_(conv(np.zeros((m + 1, k), np.float32)), A, conv(B))


@pytest.mark.parametrize("semiring", ["plus_times", "min_plus", "any_secondi", "or_and"])
def test_spmspv(env, semiring):
    gg, torch, orc = env
    print_case("spmspv", semiring)
    rng = np.random.default_rng(11)
    S = random_csr(rng, 600, 500, 0.02).tocsc()
    S.sort_indices()
    nz = np.sort(rng.choice(500, 60, replace=False)) + 1
    mask = (rng.random(600) < 0.5).astype(np.int32)
    if semiring == "plus_times":
        vals, bv, add, mul, odt = S.data, rng.random(60, dtype=np.float32), "ADD", "MUL", np.float32
    elif semiring == "min_plus":
        vals = rng.integers(1, 1000, S.nnz).astype(np.int32)
        bv, add, mul, odt = rng.integers(0, 9999, 60).astype(np.int32), "MIN", "SAFE_ADD", np.int32
    elif semiring == "any_secondi":
        vals, bv, add, mul, odt = np.ones(S.nnz, np.int32), np.ones(60, np.int32), "ANY", "SECONDI", np.int32
    else:
        vals, bv, add, mul, odt = np.ones(S.nnz, np.bool_), np.ones(60, np.bool_), "BOR", "BAND", np.bool_
    S2 = sp.csc_matrix((vals, S.indices, S.indptr), shape=S.shape)
    A = gg.SparseGPUMatrixCSC(600, 500, S.indptr + 1, S.indices + 1, vals, gg.ROCBackend(0))
    b = gg.SparseGPUVector(500, nz, bv, gg.ROCBackend(0))
    ops = {"ADD": gg.GPUGraphs_add, "MUL": gg.GPUGraphs_mul, "MIN": gg.GPUGraphs_min,
           "SAFE_ADD": gg.GPUGraphs_safe_add, "ANY": gg.GPUGraphs_any,
           "SECONDI": gg.GPUGraphs_secondi, "BOR": gg.GPUGraphs_bor, "BAND": gg.GPUGraphs_band}
    for comp in (False, True):
        out = gg.gpu_spmspv(A, b, mul=ops[mul], add=ops[add], out_dtype=odt,
                            mask=dev(torch, mask), mask_complement=comp)
        torch.cuda.synchronize()
        Ao = orc.csc(600, 500, S.indptr + 1, S.indices + 1, vals)
        ci, cv = orc.spmspv(Ao, nz, bv, 500, getattr(orc, "OP_" + add), getattr(orc, "OP_" + mul),
                            np.uint8 if odt == np.bool_ else odt, mask=mask, mask_complement=comp)
        np.testing.assert_array_equal(host(out.nzind), ci)
        gv = host(out.nzval)
        if odt == np.float32:
            np.testing.assert_allclose(gv, cv, rtol=RTOL)
        else:
            np.testing.assert_array_equal(gv.view(np.uint8) if odt == np.bool_ else gv, cv)


def _test_graphs():
    out = []
    for seed in (1, 2):
        out.append(("ba30", 30, G.barabasi_albert(30, 3, seed)))
        out.append(("dm30", 30, G.dorogovtsev_mendes(30, seed)))
    out.append(("ba2000", 2000, G.barabasi_albert(2000, 3, 5)))
    out.append(("dm3000", 3000, G.dorogovtsev_mendes(3000, 6)))
    return out


@pytest.mark.parametrize("fmt", FORMATS)
@pytest.mark.parametrize("dtype", [np.bool_, np.int32])
def test_bfs_distances_vs_oracle(env, fmt, dtype):
    """test/bfs.jl:10-46 (Bool and Int32 adjacency of BA / DM graphs)."""
    gg, torch, orc = env
    for name, n, edges in _test_graphs():
        print_case("bfs", fmt, np.dtype(dtype).name, name)
        Adj = G.undirected_adj(n, edges, dtype)
        AT = Adj.T.tocsr()
        A = make(gg, AT, fmt)
        Ao = orc_mat(orc, AT, fmt)
        for s in (1, n // 2, n):
            d = gg.bfs_distances(A, s)
            torch.cuda.synchronize()
            np.testing.assert_array_equal(host(d), orc.bfs_distances(Ao, s))


def test_bfs_directed_unreachable_and_even_values(env):
    """Directed graph with unreachable vertices; Int32 values that are even
    are not edges under or-and (2 & 1 == 0), as in the reference."""
    gg, torch, orc = env
    rng = np.random.default_rng(12)
    Adj = G.random_digraph(rng, 3000, 0.0015, wmax=4, dtype=np.int32)
    AT = Adj.T.tocsr()
    for fmt in FORMATS:
        A = make(gg, AT, fmt)   # square: no skip
        Ao = orc_mat(orc, AT, fmt)
        for s in (1, 777, 3000):
            print_case("bfs directed", fmt, s)
            np.testing.assert_array_equal(host(gg.bfs_distances(A, s)), orc.bfs_distances(Ao, s))


@pytest.mark.parametrize("fmt", FORMATS)
def test_bfs_parents_vs_oracle(env, fmt):
    """bfs_parents bit-exact including the reference's tie-break: CSR the first
    frontier in-neighbour (min col on sorted rows), SELL the largest."""
    gg, torch, orc = env
    for name, n, edges in _test_graphs():
        print_case("bfs parents", fmt, name)
        Adj = G.undirected_adj(n, edges, np.int32)
        AT = Adj.T.tocsr()
        A = make(gg, AT, fmt)
        Ao = orc_mat(orc, AT, fmt)
        for s in (1, n):
            for batch in (1, 8):
                p = gg.bfs_parents(A, s, batch_size=batch)
                torch.cuda.synchronize()
                np.testing.assert_array_equal(host(p), orc.bfs_parents(Ao, s, batch_size=batch))


def test_bfs_direction_optimising_vs_oracle(env):
    """Direction-optimising BFS levels (push over A's rows, pull over A_T)
    against the oracle's pull loop (src/algorithms/bfs.jl:31-73) on a
    symmetric R-MAT pattern in every pull layout, and on a directed graph
    whose out-edge CSR differs from A_T."""
    gg, torch, orc = env
    from gpugraphs import generators as GEN
    R = GEN.rmat(16, seed=4)
    Ao = orc.csr(R.m, R.n, R.rowptr, R.colval, np.ones(R.nnz, np.uint8))
    A = R.to_csr(gg.ROCBackend(0))
    H = R.to_sell(gg.ROCBackend(0), 64, 4096, 256, 1024)
    L = R.to_sell(gg.ROCBackend(0), 64, 1, 1024, 1024, relabel=True)
    deg = R.degrees()
    roots = list(np.nonzero(deg)[0][:5] + 1) + [int(np.argmax(deg)) + 1]
    for s in roots:
        print_case("bfs do rmat16", s)
        ref = orc.bfs_distances(Ao, int(s))
        np.testing.assert_array_equal(host(gg.bfs_distances_do(A, int(s))), ref)
        np.testing.assert_array_equal(host(gg.bfs_distances_do(H, int(s), A_out=A)), ref)
        np.testing.assert_array_equal(host(gg.bfs_distances_do(L, int(s), A_out=A)), ref)
    # directed: pushes follow A (row u = out-neighbours of u), pulls A_T
    rng = np.random.default_rng(13)
    Adj = G.random_digraph(rng, 5000, 0.002, dtype=np.bool_)
    AT = Adj.T.tocsr()
    Ad = orc.from_scipy_csr(AT)
    A_T = make(gg, AT, "csr")
    A_out = make(gg, Adj.tocsr(), "csr")
    for s in (1, 2500, 5000):
        print_case("bfs do directed", s)
        np.testing.assert_array_equal(host(gg.bfs_distances_do(A_T, s, A_out=A_out)),
                                      orc.bfs_distances(Ad, s))


@pytest.mark.parametrize("fmt", FORMATS)
def test_shortest_path_vs_oracle(env, fmt):
    """test/shortest_path.jl:31-80 (integer weights) on digraphs and a grid."""
    gg, torch, orc = env
    from gpugraphs import generators as GEN
    rng = np.random.default_rng(21)
    cases = [G.random_digraph(rng, 400, 0.01, wmax=1000, dtype=np.float32),
             G.random_digraph(rng, 400, 0.01, wmax=1000, dtype=np.int32),
             GEN.grid(40, 50, 0.01, 1, 1000, seed=2).to_scipy().T.tocsr().T]
    for Adj in cases:
        AT = sp.csr_matrix(Adj.T)
        A = make(gg, AT, fmt)
        Ao = orc_mat(orc, AT, fmt)
        for s in (1, 123, AT.shape[0]):
            print_case("sssp", fmt, AT.shape, AT.dtype, s)
            d = gg.shortest_path(A, s)
            torch.cuda.synchronize()
            np.testing.assert_array_equal(host(d), orc.shortest_path(Ao, s))


def test_shortest_path_bool_graph_hops(env):
    gg, torch, orc = env
    for name, n, edges in _test_graphs()[:2]:
        Adj = G.undirected_adj(n, edges, np.bool_)
        A = gg.SparseGPUMatrixSELL(Adj.T.tocsr(), gg.ROCBackend(0))
        d = gg.shortest_path(A, 1)
        torch.cuda.synchronize()
        assert d.dtype == torch.int32
        np.testing.assert_array_equal(host(d), host(gg.bfs_distances(A, 1)))


@pytest.mark.parametrize("fmt", FORMATS)
def test_shortest_path_multi_vs_oracle(env, fmt):
    """test/shortest_path.jl:82-126: all n sources (n x n result)."""
    gg, torch, orc = env
    rng = np.random.default_rng(4)
    Adj = G.random_digraph(rng, 60, 0.06, wmax=10, dtype=np.int32)
    AT = Adj.T.tocsr()
    A = make(gg, AT, fmt)
    srcs = np.arange(1, 61, dtype=np.int32)
    print_case("sssp multi", fmt)
    D = gg.shortest_path(A, srcs)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(host(D), orc.shortest_path_multi(orc_mat(orc, AT, fmt), srcs))


@pytest.mark.parametrize("k", [1, 64, 100, 256])
def test_multi_source_bfs_equals_single(env, k):
    gg, torch, orc = env
    from gpugraphs import generators as GEN
    R = GEN.rmat(13, seed=9)
    rng = np.random.default_rng(k)
    nzr = np.nonzero(R.degrees())[0] + 1
    srcs = rng.choice(nzr, k, replace=True).astype(np.int32)
    for A in (R.to_csr(gg.ROCBackend(0)), R.to_sell(gg.ROCBackend(0), 64, 4096, 128, 256),
              R.to_sell(gg.ROCBackend(0), 64, 1, 128, 256, relabel=True)):
        print_case("msbfs", k, type(A).__name__, A.vinv is not None if hasattr(A, "vinv") else "")
        D = host(gg.multi_source_bfs(A, srcs))
        Ao = orc.from_scipy_csr(R.to_scipy())
        for i in rng.choice(k, min(k, 6), replace=False):
            np.testing.assert_array_equal(D[:, i], orc.bfs_distances(Ao, int(srcs[i])))


@pytest.mark.parametrize("fmt", ["csr", "sell_relabel"])
def test_multi_source_bfs_deep_levels(env, fmt):
    """More than 254 levels (a path of 700 vertices): the byte-level scratch
    overflows and the batch is traversed again with Int32 levels."""
    gg, torch, orc = env
    n = 700
    S = sp.diags([np.ones(n - 1), np.ones(n - 1)], [-1, 1], format="csr").astype(np.int32)
    A = make(gg, S, fmt)
    srcs = np.array([1, n, n // 2, 3], np.int32)
    print_case("msbfs deep", fmt)
    D = host(gg.multi_source_bfs(A, srcs))
    Ao = orc.from_scipy_csr(S)
    for i, s in enumerate(srcs):
        np.testing.assert_array_equal(D[:, i], orc.bfs_distances(Ao, int(s)))
    assert D.max() == n - 1


@pytest.mark.parametrize("steps", [0, 2, 6, 32])
@pytest.mark.parametrize("fmt", ["csr", "sell_relabel"])
def test_multi_source_bfs_level_windows(env, steps, fmt):
    """Byte levels recorded a window of levels at a time (every 8 host-stepped
    levels; captured: every 2, 6 or 16 levels): a path of 230 vertices, a
    20 x 20 grid and an R-MAT graph as disjoint components, 100 sources, so
    that one group of 8 sources is reached in many windows and most vertices
    stay unreached by most sources (typemax(Int32))."""
    gg, torch, orc = env
    from gpugraphs import _lib, generators as GEN
    n1 = 230
    path = sp.diags([np.ones(n1 - 1), np.ones(n1 - 1)], [-1, 1], format="csr")
    g = sp.diags([np.ones(399), np.ones(399)], [-1, 1], format="lil")
    for i in range(19, 399, 20):
        g[i, i + 1] = 0
        g[i + 1, i] = 0
    grid = (sp.csr_matrix(g) + sp.diags([np.ones(380), np.ones(380)], [-20, 20], format="csr"))
    grid.eliminate_zeros()
    R = GEN.rmat(10, seed=8).to_scipy()
    S = sp.block_diag([path, grid, R], format="csr").astype(np.int32)
    S.data[:] = 1
    rng = np.random.default_rng(steps + 3)
    srcs = rng.choice(S.shape[0], 100, replace=False).astype(np.int32) + 1
    srcs[:4] = [1, n1, n1 // 2, n1 + 1]
    A = make(gg, S, fmt)
    Ao = orc.from_scipy_csr(S)
    ctx = _lib.context(0)
    try:
        _lib.set_graph_steps(ctx, steps)
        print_case("msbfs windows", steps, fmt)
        D = host(gg.multi_source_bfs(A, srcs))
    finally:
        _lib.set_graph_steps(ctx, 0)
    for i, s in enumerate(srcs):
        np.testing.assert_array_equal(D[:, i], orc.bfs_distances(Ao, int(s)))
    assert D[:n1, 0].max() == n1 - 1


def _host_csr(S, dtype):
    from gpugraphs import generators as GEN
    S = sp.csr_matrix(S)
    S.sort_indices()
    return GEN.HostCSR(S.shape[0], S.shape[1], (S.indptr + 1).astype(np.int32),
                       (S.indices + 1).astype(np.int32), S.data.astype(dtype))


@pytest.mark.parametrize("dtype", [np.bool_, np.int32])
def test_bfs_packed_equals_per_graph(env, dtype):
    """One persistent launch over 40 block-diagonally packed graphs of uneven
    sizes (1 to ~2K vertices) equals bfs_distances per graph; Int32 values
    that are even are not edges, as in the single-graph BFS."""
    gg, torch, orc = env
    from gpugraphs import generators as GEN
    rng = np.random.default_rng(77 if dtype == np.bool_ else 78)
    mats = [sp.csr_matrix(np.ones((1, 1)) - np.eye(1))]            # one isolated vertex
    for i in range(39):
        kind = i % 4
        if kind == 0:
            n = int(rng.integers(10, 600))
            M = G.undirected_adj(n, G.barabasi_albert(n, 3, int(rng.integers(1 << 30))), np.int32)
        elif kind == 1:
            M = G.random_digraph(rng, int(rng.integers(5, 400)), 0.02, wmax=4, dtype=np.int32)
        elif kind == 2:
            M = GEN.rmat(int(rng.integers(4, 12)), seed=int(rng.integers(1 << 30))).to_scipy()
        else:
            n = int(rng.integers(2, 300))
            M = G.undirected_adj(n, G.dorogovtsev_mendes(n, int(rng.integers(1 << 30))), np.int32)
        if dtype == np.bool_:
            M = (M != 0).astype(np.int32)
        mats.append(sp.csr_matrix(M.T))
    hosts = [_host_csr(M, dtype) for M in mats]
    P, off = GEN.pack(hosts)
    A = P.to_csr(gg.ROCBackend(0))
    srcs = np.array([int(rng.integers(M.shape[0])) + 1 for M in mats], np.int32)
    print_case("bfs packed", np.dtype(dtype).name, P.m, P.nnz)
    D, levels = gg.bfs_packed(A, off, srcs)
    D = host(D)
    for g, M in enumerate(mats):
        Ao = orc.from_scipy_csr(sp.csr_matrix(M.astype(np.int32)) if dtype == np.int32 else
                                sp.csr_matrix(M.astype(np.bool_)))
        ref = orc.bfs_distances(Ao, int(srcs[g]))
        blk = D[off[g]:off[g] + M.shape[0]]
        np.testing.assert_array_equal(blk, ref)
        assert np.all(D[off[g] + M.shape[0]:off[g + 1]] == INT32_MAX)   # padding unreached
        fin = ref[ref != INT32_MAX]
        assert levels[g] == fin.max()
    # again, after other work on the context
    gg.bfs_distances(A, 1)
    D2, _ = gg.bfs_packed(A, off, srcs)
    np.testing.assert_array_equal(host(D2), D)


def _sym_host(S):
    from gpugraphs import generators as GEN
    S = sp.csr_matrix(S)
    S.sort_indices()
    assert (S != S.T).nnz == 0
    return GEN.HostCSR(S.shape[0], S.shape[1], (S.indptr + 1).astype(np.int32),
                       (S.indices + 1).astype(np.int32), np.ones(S.nnz, np.bool_), True)


@pytest.mark.parametrize("steps,alpha,k", [(0, 16.0, 20), (8, 16.0, 20), (0, 2.0, 100),
                                           (8, 2.0, 20), (0, 1e-9, 20), (8, 1e-9, 100)])
def test_multi_source_bfs_top_down(env, steps, alpha, k):
    """Symmetric unit patterns switch levels with small frontiers to the
    top-down step (rows longer than 1024 entries cut into pieces; a short
    queue spread over many blocks; a wide push clears the next front so that
    a dense pull may follow; alpha 1e-9 pushes every level); one and two mask
    words per vertex.  The levels equal the single-source BFS."""
    gg, torch, orc = env
    from gpugraphs import _lib, generators as GEN
    n = 12001
    star = sp.coo_matrix((np.ones(2 * (n - 1)), (np.r_[np.zeros(n - 1), np.arange(1, n)],
                                                 np.r_[np.arange(1, n), np.zeros(n - 1)])),
                         shape=(n, n)).tocsr()
    path = sp.diags([np.ones(799), np.ones(799)], [-1, 1], format="csr")
    R = GEN.rmat(13, seed=11)
    cases = [("star", _sym_host(star)), ("path", _sym_host(path)), ("rmat13", R)]
    ctx = _lib.context(0)
    try:
        _lib.set_graph_steps(ctx, steps)
        _lib.set_push_threshold(ctx, alpha)
        for name, H in cases:
            assert H.symmetric
            Ao = orc.from_scipy_csr(H.to_scipy())
            rng = np.random.default_rng(len(name))
            srcs = rng.choice(np.nonzero(H.degrees())[0], k, replace=False).astype(np.int32) + 1
            srcs[0] = 1
            want = [orc.bfs_distances(Ao, int(s)) for s in srcs]
            for A in (H.to_csr(gg.ROCBackend(0)),
                      H.to_sell(gg.ROCBackend(0), 64, 1, 256, 512, relabel=True),
                      H.to_sell(gg.ROCBackend(0), 64, 1, 256, 512, relabel=True,
                                sort_entries=True)):
                print_case("msbfs top-down", steps, alpha, k, name, type(A).__name__)
                D = host(gg.multi_source_bfs(A, srcs))
                for i in range(len(srcs)):
                    np.testing.assert_array_equal(D[:, i], want[i])
    finally:
        _lib.set_graph_steps(ctx, 0)
        _lib.set_push_threshold(ctx, 16.0)


@pytest.mark.parametrize("steps", [0, 8])
def test_shortest_path_sparse_iterations(env, steps):
    """Symmetric patterns (independent weights per direction) on a relabelled
    SELL: k-source iterations relax only the rows next to the rows that fell
    when those are few; distances equal the oracle's, one and many sources,
    host-stepped and captured."""
    gg, torch, orc = env
    from gpugraphs import _lib, generators as GEN
    rng = np.random.default_rng(5)
    R = GEN.rmat(12, seed=4)
    w = rng.integers(1, 100, R.nnz).astype(np.int32)
    cases = [GEN.grid(60, 70, 0.01, 1, 1000, seed=9, dtype=np.float32),
             GEN.grid(33, 41, 0.0, 1, 50, seed=2, dtype=np.int32),
             GEN.HostCSR(R.m, R.n, R.rowptr, R.colval, w, True)]
    ctx = _lib.context(0)
    try:
        _lib.set_graph_steps(ctx, steps)
        for H in cases:
            assert H.symmetric
            S = H.to_scipy()
            Ao = orc.from_scipy_csr(S)
            A = H.to_sell(gg.ROCBackend(0), 64, 1, 256, 256, relabel=True)
            srcs = (rng.choice(np.nonzero(H.degrees())[0], 6, replace=False) + 1).astype(np.int32)
            print_case("sssp sparse", steps, H.m, H.nzval.dtype)
            for s in srcs[:3]:
                np.testing.assert_array_equal(host(gg.shortest_path(A, int(s))),
                                              orc.shortest_path(Ao, int(s)))
            np.testing.assert_array_equal(host(gg.shortest_path(A, srcs)),
                                          orc.shortest_path_multi(Ao, srcs))
    finally:
        _lib.set_graph_steps(ctx, 0)


@pytest.mark.parametrize("dt", [np.int32, np.float32, np.int64, np.float64])
@pytest.mark.parametrize("k", [2, 8, 12, 64, 72])
def test_shortest_path_multi_column_vectors(env, dt, k):
    """k sources with 16-byte column groups per lane (k a multiple of 4 for
    32-bit distances, of 2 for 64-bit ones) on SELL layouts with heavy rows,
    and the one-column-per-lane path otherwise: distances equal the oracle's."""
    gg, torch, orc = env
    from gpugraphs import generators as GEN
    rng = np.random.default_rng(k * 7 + np.dtype(dt).itemsize)
    R = GEN.rmat(11, seed=k)
    w = rng.integers(1, 200, R.nnz).astype(dt)
    H = GEN.HostCSR(R.m, R.n, R.rowptr, R.colval, w, True)
    Ao = orc.from_scipy_csr(H.to_scipy())
    srcs = (rng.choice(np.nonzero(H.degrees())[0], k, replace=True) + 1).astype(np.int32)
    ref = orc.shortest_path_multi(Ao, srcs)
    for A in (H.to_sell(gg.ROCBackend(0), 64, 1, 64, 128, relabel=True),
              H.to_sell(gg.ROCBackend(0), 64, 1, 32, 2048, relabel=True),
              H.to_sell(gg.ROCBackend(0), 64, 256, 64, 128)):
        print_case("sssp column vectors", np.dtype(dt).name, k, A.vinv is not None)
        np.testing.assert_array_equal(host(gg.shortest_path(A, srcs)), ref)


def _sell_arrays(torch, M):
    def h(x):
        if x is None:
            return None
        return x.cpu().numpy() if torch.is_tensor(x) else np.asarray(x)
    out = {k: h(getattr(M, k)) for k in ("slice_ptr", "colval", "nzval", "slot_len", "heavy_row",
                                         "heavy_ptr", "heavy_seg", "heavy_col", "heavy_val")}
    out["perm_dev"] = h(M._perm_dev)
    out["vinv"] = h(M.vinv)
    out["scalars"] = (M.m, M.n, M.nnz(), M.slice_size, M.nslices, M.n_stored, M.n_heavy, M.n_seg)
    return out


@pytest.mark.parametrize("layout", [(32, 1, 0, 4096, False), (64, 64, 8, 16, False),
                                    (64, 1, 8, 16, True), (64, 4096, 256, 512, False),
                                    (16, 1, 0, 4096, True), (64, 1, 0, 4096, False),
                                    (8, 3, 5, 7, False)])
@pytest.mark.parametrize("sort_entries", [False, True])
def test_sell_device_builder_equals_host(env, layout, sort_entries):
    """The device builder (gg_sell_plan_device / gg_sell_fill_device) gives
    the host builder's arrays exactly, for the reference layout, sigma
    windows, heavy partitions and relabelling, and the same SpMV."""
    gg, torch, orc = env
    from gpugraphs import generators as GEN
    C_, sigma, heavy, seg, relabel = layout
    rng = np.random.default_rng(sum(layout[:4]))
    cases = [GEN.rmat(11, seed=2), GEN.grid(30, 40, 0.02, 1, 50, seed=3, dtype=np.float32)]
    S = random_csr(rng, 300, 300, 0.03).astype(np.int32)
    S.data[:] = rng.integers(1, 9, S.nnz)
    S.data[S.indptr[5]:S.indptr[6]] = 0   # an empty row
    S.eliminate_zeros()
    cases.append(GEN.HostCSR(300, 300, (S.indptr + 1).astype(np.int32),
                             (S.indices + 1).astype(np.int32), S.data.astype(np.int32)))
    if not relabel:
        R = random_csr(rng, 200, 500, 0.02).astype(np.float64)
        cases.append(GEN.HostCSR(200, 500, (R.indptr + 1).astype(np.int32),
                                 (R.indices + 1).astype(np.int32), R.data))
    for H in cases:
        print_case("sell device builder", layout, H.m, H.n, H.nzval.dtype)
        A = H.to_csr(gg.ROCBackend(0))
        ref = H.to_sell(gg.ROCBackend(0), C_, sigma, heavy, seg, relabel, sort_entries)
        got = gg.SparseGPUMatrixSELL.build_on_device(A, C_, sigma, heavy, seg, relabel,
                                                     sort_entries)
        a, b = _sell_arrays(torch, ref), _sell_arrays(torch, got)
        for k in a:
            if a[k] is None or b[k] is None:
                assert a[k] is None and b[k] is None or (a[k] is not None and len(a[k]) == 0) \
                    or (b[k] is not None and len(b[k]) == 0), k
            else:
                np.testing.assert_array_equal(np.asarray(a[k]), np.asarray(b[k]), err_msg=k)
        x = dev(torch, rng.random(H.n).astype(np.float32) if H.nzval.dtype == np.float32
                else rng.integers(0, 5, H.n).astype(H.nzval.dtype if H.nzval.dtype != np.bool_
                                                    else np.int32))
        if H.nzval.dtype == np.bool_:
            continue
        y1 = dev(torch, np.zeros(H.m, H.nzval.dtype))
        y2 = dev(torch, np.zeros(H.m, H.nzval.dtype))
        gg.gpu_spmv_(y1, ref, x)
        gg.gpu_spmv_(y2, got, x)
        np.testing.assert_array_equal(host(y1), host(y2))


@pytest.mark.parametrize("relabel", [False, True])
def test_spmm_power_law_heavy_rows(env, relabel):
    """SpMM on a power-law matrix whose hubs span several heavy segments:
    plus-times within the tolerance of an FP64 product, the min-plus forms
    exact against the oracle (16-byte column groups, segment partials folded
    in order)."""
    gg, torch, orc = env
    from gpugraphs import generators as GEN
    rng = np.random.default_rng(31 + relabel)
    R = GEN.rmat(13, seed=5, dtype=np.float32, values=GEN.UNIFORM01)
    S = R.to_scipy()
    A = R.to_sell(gg.ROCBackend(0), 64, 1, 64, 256, relabel=relabel)
    k = 32
    B = rng.random((R.n, k), dtype=np.float32)
    C = dev(torch, np.zeros((R.m, k), np.float32))
    print_case("spmm power law", relabel)
    gg.gpu_spmm_(C, A, dev(torch, B))
    np.testing.assert_allclose(host(C), S.astype(np.float64) @ B.astype(np.float64), rtol=RTOL,
                               atol=1e-6)
    W = R.to_scipy()
    W.data = rng.integers(1, 500, W.nnz).astype(np.float32)
    Hw = GEN.HostCSR(R.m, R.n, R.rowptr, R.colval, W.data.astype(np.float32), True)
    Aw = Hw.to_sell(gg.ROCBackend(0), 64, 1, 64, 256, relabel=relabel)
    Bi = rng.integers(0, 3000, (R.n, k)).astype(np.float32)
    Bi[rng.random((R.n, k)) < 0.2] = np.inf
    C0 = rng.integers(0, 9000, (R.m, k)).astype(np.float32)
    for add, mul, acc in (("min", "safe_add", "min"), ("min", "add", "second")):
        Cd = dev(torch, C0.copy())
        gg.gpu_spmm_(Cd, Aw, dev(torch, Bi), mul=_ops(gg, mul), add=_ops(gg, add),
                     accum=_ops(gg, acc))
        Co = np.asfortranarray(C0.copy())
        orc.spmm(Co, orc.from_scipy_csr(Hw.to_scipy()), np.asfortranarray(Bi),
                 add=getattr(orc, "OP_" + add.upper()), mul=getattr(orc, "OP_" + mul.upper()),
                 accum=getattr(orc, "OP_" + acc.upper()))
        np.testing.assert_array_equal(host(Cd), Co)


def test_sell_device_builder_refuses_rcm(env):
    """Reverse Cuthill-McKee ties are a host-builder layout: the device builder
    rejects the flag before any launch, and the context keeps working."""
    gg, torch, orc = env
    from gpugraphs import generators as GEN
    R = GEN.rmat(8, seed=2)
    A = R.to_csr(gg.ROCBackend(0))
    print_case("sell device builder refuses rcm")
    with pytest.raises(gg.ArgumentError):
        gg.SparseGPUMatrixSELL.build_on_device(A, 64, 1, 0, 4096, True, True, rcm=True)
    got = gg.SparseGPUMatrixSELL.build_on_device(A, 64, 1, 0, 4096, True, True)
    Ao = orc.from_scipy_csr(R.to_scipy())
    s = int(np.nonzero(R.degrees())[0][0]) + 1
    np.testing.assert_array_equal(host(gg.bfs_distances(got, s)), orc.bfs_distances(Ao, s))


@pytest.mark.parametrize("relabel", [False, True])
def test_sell_device_builder_degenerate(env, relabel):
    """Rows without entries and a single-row matrix: the device builder still
    equals the host builder."""
    gg, torch, orc = env
    from gpugraphs import generators as GEN
    cases = [GEN.HostCSR(7, 7, np.ones(8, np.int32), np.zeros(0, np.int32),
                         np.zeros(0, np.float32)),
             GEN.HostCSR(1, 1, np.array([1, 2], np.int32), np.array([1], np.int32),
                         np.array([2.5], np.float32))]
    for H in cases:
        for layout in ((32, 1, 0, 4096), (64, 8, 1, 4)):
            print_case("sell device builder degenerate", H.m, layout, relabel)
            ref = H.to_sell(gg.ROCBackend(0), *layout, relabel=relabel)
            got = gg.SparseGPUMatrixSELL.build_on_device(H.to_csr(gg.ROCBackend(0)), *layout,
                                                         relabel=relabel)
            a, b = _sell_arrays(torch, ref), _sell_arrays(torch, got)
            for key in a:
                if a[key] is None or b[key] is None:
                    assert (a[key] is None or len(a[key]) == 0) and \
                        (b[key] is None or len(b[key]) == 0), key
                else:
                    np.testing.assert_array_equal(a[key], b[key], err_msg=key)


def test_bfs_packed_rejects_unpadded_last_graph(env):
    """The persistent kernel walks whole 256-row chunks: a pack whose last
    block is not padded to a multiple of 256 rows is an ArgumentError (its
    trailing rows would otherwise stay unvisited)."""
    gg, torch, orc = env
    from gpugraphs import generators as GEN
    rng = np.random.default_rng(79)
    M = G.undirected_adj(300, G.barabasi_albert(300, 3, 5), np.int32).astype(np.bool_)
    P, off = GEN.pack([_host_csr(sp.csr_matrix(M.T), np.bool_)])
    assert P.m == 512 and off.tolist() == [0, 512]
    # drop the padding rows of the only block: 300 rows, offsets [0, 300]
    Hs = GEN.HostCSR(300, 300, P.rowptr[:301].copy(), P.colval, P.nzval, True)
    A = Hs.to_csr(gg.ROCBackend(0))
    print_case("bfs packed unpadded")
    with pytest.raises(gg.ArgumentError):
        gg.bfs_packed(A, np.array([0, 300], np.int64), np.array([int(rng.integers(300)) + 1],
                                                                np.int32))


@pytest.mark.parametrize("alpha", [1e-9, 16.0, 0.0])
def test_bfs_packed_symmetric_vs_oracle(env, alpha):
    """Symmetric unit-pattern packs (Barabasi-Albert, Dorogovtsev-Mendes,
    R-MAT and grid blocks, one source a hub) start with top-down levels while
    the queue's degree sum * alpha is below the entries left (alpha 1e-9:
    every level top-down until all graphs finish; 16: top-down, then the
    persistent pull kernel; 0: pull only); every graph's levels equal the
    oracle's pull loop (src/algorithms/bfs.jl:31-73) from its source."""
    gg, torch, orc = env
    from gpugraphs import _lib
    from gpugraphs import generators as GEN
    rng = np.random.default_rng(91)
    hosts, mats = [], []
    for i in range(48):
        kind = i % 4
        if kind == 0:
            n = int(rng.integers(10, 900))
            M = G.undirected_adj(n, G.barabasi_albert(n, 3, int(rng.integers(1 << 30))), np.bool_)
        elif kind == 1:
            n = int(rng.integers(2, 400))
            M = G.undirected_adj(n, G.dorogovtsev_mendes(n, int(rng.integers(1 << 30))), np.bool_)
        elif kind == 2:
            M = GEN.rmat(int(rng.integers(4, 13)), seed=int(rng.integers(1 << 30))).to_scipy()
        else:
            M = GEN.grid(int(rng.integers(2, 40)), int(rng.integers(2, 40)), 0.0, 1, 1,
                         seed=int(rng.integers(1 << 30)), dtype=np.int32).to_scipy()
        M = sp.csr_matrix(M != 0).astype(np.bool_)
        H = _host_csr(M, np.bool_)
        H.symmetric = True
        hosts.append(H)
        mats.append(M)
    P, off = GEN.pack(hosts)
    assert P.symmetric
    A = P.to_csr(gg.ROCBackend(0))
    srcs = np.array([int(rng.integers(M.shape[0])) + 1 for M in mats], np.int32)
    srcs[0] = int(np.argmax(np.diff(mats[0].indptr))) + 1       # a hub
    ctx = _lib.context(0)
    try:
        _lib.set_push_threshold(ctx, alpha)
        print_case("bfs packed symmetric", alpha, P.m, P.nnz)
        D, levels = gg.bfs_packed(A, off, srcs)
    finally:
        _lib.set_push_threshold(ctx, 16.0)
    D = host(D)
    for g, M in enumerate(mats):
        ref = orc.bfs_distances(orc.from_scipy_csr(sp.csr_matrix(M)), int(srcs[g]))
        np.testing.assert_array_equal(D[off[g]:off[g] + M.shape[0]], ref, err_msg=f"graph {g}")
        assert np.all(D[off[g] + M.shape[0]:off[g + 1]] == INT32_MAX)
        assert levels[g] == ref[ref != INT32_MAX].max()


def _unpadded_pack(hosts):
    """Block-diagonal pack of A_T blocks with no padding rows: (HostCSR, offsets)."""
    from gpugraphs import generators as GEN
    blocks = [sp.csr_matrix((h.nzval, h.colval - 1, h.rowptr.astype(np.int64) - 1),
                            shape=(h.m, h.n)) for h in hosts]
    S = sp.block_diag(blocks, format="csr").astype(np.bool_)
    S.sort_indices()
    off = np.r_[0, np.cumsum([h.m for h in hosts])].astype(np.int64)
    return GEN.HostCSR(S.shape[0], S.shape[1], (S.indptr + 1).astype(np.int32),
                       (S.indices + 1).astype(np.int32), np.ones(S.nnz, np.bool_),
                       all(h.symmetric for h in hosts)), off


@pytest.mark.parametrize("alpha", [1e-9, 15.0, 1e9])
@pytest.mark.parametrize("layout", ["csr", "sell_relabel", "sell_sigma"])
def test_bfs_packed_do_vs_oracle(env, layout, alpha):
    """One direction-optimising traversal of a pack seeded with every graph's
    source (gg_bfs_packed_do) gives each graph the oracle's levels
    (src/algorithms/bfs.jl:31-73) from its source: symmetric blocks (R-MAT,
    Barabasi-Albert, grids, a lone vertex, a hub source) and directed blocks
    whose out-edge CSR differs from A_T, padded (gg_pack_graphs) and unpadded
    offsets; alpha 1e-9 pushes every level, 1e9 pulls every level; the pull
    runs over CSR, a relabelled SELL with heavy rows, and sigma windows."""
    gg, torch, orc = env
    from gpugraphs import generators as GEN
    rng = np.random.default_rng(123)
    sym_hosts, dir_hosts, sym_mats, dir_mats = [], [], [], []
    sym_mats.append(sp.csr_matrix((1, 1), dtype=np.bool_))               # a lone vertex
    for i in range(23):
        kind = i % 3
        if kind == 0:
            M = GEN.rmat(int(rng.integers(4, 13)), seed=int(rng.integers(1 << 30))).to_scipy()
        elif kind == 1:
            n = int(rng.integers(10, 900))
            M = G.undirected_adj(n, G.barabasi_albert(n, 3, int(rng.integers(1 << 30))), np.bool_)
        else:
            M = GEN.grid(int(rng.integers(2, 40)), int(rng.integers(2, 40)), 0.0, 1, 1,
                         seed=int(rng.integers(1 << 30)), dtype=np.int32).to_scipy()
        sym_mats.append(sp.csr_matrix(M != 0).astype(np.bool_))
    for M in sym_mats:
        H = _host_csr(M, np.bool_)
        H.symmetric = True
        sym_hosts.append(H)
    for i in range(12):
        Adj = G.random_digraph(rng, int(rng.integers(5, 700)), 0.01, dtype=np.bool_)
        AT = sp.csr_matrix(Adj.T).astype(np.bool_)
        dir_mats.append(AT)
        dir_hosts.append(_host_csr(AT, np.bool_))
    R = gg.ROCBackend(0)
    for name, hosts, mats in (("symmetric", sym_hosts, sym_mats),
                              ("directed", dir_hosts, dir_mats)):
        for padded in (True, False):
            P, off = GEN.pack(hosts) if padded else _unpadded_pack(hosts)
            A_T_csr = P.to_csr(R)
            if name == "symmetric":
                A_out = A_T_csr
            else:
                A_out = gg.SparseGPUMatrixCSR(sp.csr_matrix(P.to_scipy().T), R)
            if layout == "csr":
                A_T = A_T_csr
            elif layout == "sell_relabel":
                A_T = P.to_sell(R, 64, 1, 8, 16, relabel=True, sort_entries=True)
            else:
                A_T = P.to_sell(R, 64, 64, 8, 16)
            srcs = np.array([int(rng.integers(M.shape[0])) + 1 for M in mats], np.int32)
            if name == "symmetric":
                srcs[1] = int(np.argmax(np.diff(mats[1].indptr))) + 1   # a hub
            print_case("bfs packed do", name, layout, alpha, padded, P.m, P.nnz)
            D, levels = gg.bfs_packed_do(A_T, off, srcs, A_out=A_out, alpha=alpha)
            D = host(D)
            for g, M in enumerate(mats):
                ref = orc.bfs_distances(orc.from_scipy_csr(sp.csr_matrix(M)), int(srcs[g]))
                np.testing.assert_array_equal(D[off[g]:off[g] + M.shape[0]], ref,
                                              err_msg=f"{name} graph {g}")
                assert np.all(D[off[g] + M.shape[0]:off[g + 1]] == INT32_MAX)
                assert levels[g] == ref[ref != INT32_MAX].max(), f"{name} graph {g}"
            if padded and layout == "csr":
                # the persistent launch gives the same arrays
                D2, lv2 = gg.bfs_packed(A_T_csr, off, srcs)
                np.testing.assert_array_equal(host(D2), D)
                np.testing.assert_array_equal(lv2, levels)


def test_bfs_packed_do_errors(env):
    """Offsets that do not cover the rows, a source outside its graph, a
    missing or non-CSR A_out are rejected before any launch."""
    gg, torch, orc = env
    from gpugraphs import generators as GEN
    hosts = [GEN.rmat(6, seed=s) for s in (1, 2)]
    P, off = GEN.pack(hosts)
    A = P.to_csr(gg.ROCBackend(0))
    S = P.to_sell(gg.ROCBackend(0), 64, 1, 0, 4096, relabel=True)
    srcs = np.array([1, 1], np.int32)
    print_case("bfs packed do errors")
    with pytest.raises(gg.DimensionMismatch):
        gg.bfs_packed_do(A, off[:-1].tolist() + [off[-1] - 1], srcs)
    with pytest.raises(gg.BoundsError):
        gg.bfs_packed_do(A, off, np.array([1, off[2] - off[1] + 1], np.int32))
    with pytest.raises(gg.ArgumentError):
        gg.bfs_packed_do(S, off, srcs)
    with pytest.raises(gg.ArgumentError):
        gg.bfs_packed_do(S, off, srcs, A_out=S)
    D, lv = gg.bfs_packed_do(S, off, srcs, A_out=A)
    D2, lv2 = gg.bfs_packed(A, off, srcs)
    np.testing.assert_array_equal(host(D), host(D2))
    np.testing.assert_array_equal(lv, lv2)


@pytest.mark.parametrize("steps", [0, 8])
def test_shortest_path_delta_vs_oracle(env, steps):
    """Delta-stepping (gg_shortest_path_delta, push over the CSR of A) gives
    the distances of the oracle's Jacobi Bellman-Ford
    (src/algorithms/shortest_path.jl:34-83) from A_T: weighted digraphs
    (fp32, Int32 with zero weights, Int64, fp64), a grid with long-range edges
    and independent weights per direction, a weighted R-MAT whose hubs are cut
    into pieces, unreachable vertices, Bool hops; delta from one (a bucket per
    distance) through the default to 1e12 (one bucket: push Bellman-Ford);
    host-stepped and captured phases."""
    gg, torch, orc = env
    from gpugraphs import _lib, generators as GEN
    rng = np.random.default_rng(31)
    R = GEN.rmat(14, seed=8)
    w = rng.integers(1, 1000, R.nnz).astype(np.float32)
    RW = sp.csr_matrix((w, R.colval - 1, R.rowptr.astype(np.int64) - 1), shape=(R.m, R.n))
    Z = G.random_digraph(rng, 500, 0.01, wmax=4, dtype=np.int32)
    Z.data -= 1                                                     # weights 0..3
    cases = [("digraph f32", G.random_digraph(rng, 600, 0.008, wmax=1000, dtype=np.float32)),
             ("digraph i32 zero weights", Z),
             ("digraph i64", G.random_digraph(rng, 300, 0.02, wmax=50, dtype=np.int64)),
             ("digraph f64", G.random_digraph(rng, 300, 0.02, wmax=50, dtype=np.float64)),
             ("grid", GEN.grid(60, 70, 0.01, 1, 1000, seed=9, dtype=np.float32).to_scipy().T.tocsr()),
             ("rmat14 weighted", RW.T.tocsr()),
             ("sparse digraph", G.random_digraph(rng, 400, 0.002, wmax=9, dtype=np.float32))]
    ctx = _lib.context(0)
    try:
        _lib.set_graph_steps(ctx, steps)
        for name, Adj in cases:
            Adj = sp.csr_matrix(Adj)
            Adj.sort_indices()
            AT = sp.csr_matrix(Adj.T)
            Ao = orc.from_scipy_csr(AT)
            A_out = gg.SparseGPUMatrixCSR(Adj, gg.ROCBackend(0))
            deg = np.diff(Adj.indptr)
            srcs = [1, int(np.argmax(deg)) + 1, Adj.shape[0]]
            for s in srcs:
                ref = orc.shortest_path(Ao, s)
                for delta in (1.0, 0.0, 1e12):
                    print_case("sssp delta", name, steps, s, delta)
                    d = gg.shortest_path_delta(A_out, s, delta)
                    assert d.dtype == torch.from_numpy(ref[:1]).dtype
                    np.testing.assert_array_equal(host(d), ref, err_msg=f"{name} {s} {delta}")
            # k sources sharing their phases (a repeated source included)
            ks = np.array(srcs + [int(x) for x in rng.integers(1, Adj.shape[0] + 1, 6)] + [srcs[0]],
                          np.int32)
            ref_k = orc.shortest_path_multi(Ao, ks)
            for delta in (1.0, 0.0, 1e12):
                print_case("sssp delta multi", name, steps, len(ks), delta)
                D = gg.shortest_path_delta(A_out, ks, delta)
                assert D.shape == (Adj.shape[0], len(ks))
                np.testing.assert_array_equal(host(D), ref_k, err_msg=f"{name} k {delta}")
        # Bool: hop counts as Int32
        Adj = sp.csr_matrix(G.random_digraph(rng, 800, 0.004, dtype=np.bool_))
        A_out = gg.SparseGPUMatrixCSR(Adj, gg.ROCBackend(0))
        Ao = orc.from_scipy_csr(sp.csr_matrix(Adj.T))
        for s in (1, 400):
            print_case("sssp delta bool", steps, s)
            d = gg.shortest_path_delta(A_out, s)
            assert d.dtype == torch.int32
            np.testing.assert_array_equal(host(d), orc.shortest_path(Ao, s))
    finally:
        _lib.set_graph_steps(ctx, 0)


def test_shortest_path_delta_errors(env):
    """A negative weight reached from a source is an ArgumentError (one that
    is not reached does not matter); a source outside 1:n is a BoundsError; a
    SELL matrix, a negative delta or more than 1024 sources is an
    ArgumentError."""
    gg, torch, orc = env
    rng = np.random.default_rng(32)
    Adj = sp.csr_matrix(G.random_digraph(rng, 200, 0.03, wmax=9, dtype=np.float32))
    neg = Adj.copy()
    neg.data[0] = -1.0
    r0 = int(np.searchsorted(neg.indptr, 0, side="right")) - 1     # row of entry 0
    A_neg = gg.SparseGPUMatrixCSR(neg, gg.ROCBackend(0))
    print_case("sssp delta errors")
    with pytest.raises(gg.ArgumentError):
        gg.shortest_path_delta(A_neg, r0 + 1)
    # a negative edge out of an unreachable vertex: vertex 201 has no in-edges
    iso = sp.bmat([[Adj, None], [sp.csr_matrix(np.r_[-5.0, np.zeros(199)][None, :]), None]],
                  format="csr").astype(np.float32)
    iso.resize((201, 201))
    iso.eliminate_zeros()
    A_iso = gg.SparseGPUMatrixCSR(sp.csr_matrix(iso), gg.ROCBackend(0))
    Ao = orc.from_scipy_csr(sp.csr_matrix(iso.T))
    np.testing.assert_array_equal(host(gg.shortest_path_delta(A_iso, 1)), orc.shortest_path(Ao, 1))
    A = gg.SparseGPUMatrixCSR(Adj, gg.ROCBackend(0))
    with pytest.raises(gg.BoundsError):
        gg.shortest_path_delta(A, 201)
    with pytest.raises(gg.ArgumentError):
        gg.shortest_path_delta(A, 1, -1.0)
    with pytest.raises(gg.ArgumentError):
        gg.shortest_path_delta(gg.SparseGPUMatrixSELL(Adj, gg.ROCBackend(0), 32), 1)
    with pytest.raises(gg.BoundsError):
        gg.shortest_path_delta(A, [1, 201])
    with pytest.raises(gg.ArgumentError):
        gg.shortest_path_delta(A, np.ones(1025, np.int32))
    with pytest.raises(gg.ArgumentError):
        gg.shortest_path_delta(A_neg, [1, r0 + 1])
    # the context still works after the errors
    np.testing.assert_array_equal(host(gg.shortest_path_delta(A, 1)),
                                  orc.shortest_path(orc.from_scipy_csr(sp.csr_matrix(Adj.T)), 1))
