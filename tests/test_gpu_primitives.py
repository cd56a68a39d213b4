"""GPU parity of the L1 primitives behind the C ABI: zgemm (the MKL call sites of
Recursive_Greens.F90 / Contact_Self_Energy.F90) and Invert_Matrix
(MKL_Routines_Cluster.F90:2-33), checked against numpy/LAPACK complex128."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

capi = pytest.importorskip("nonequilibriumgreensfunction_amd.capi")


def _rand(rng, n):
    return (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))).astype(np.complex128)


def _op(t, M):
    return {"N": M, "T": M.T, "C": M.conj().T}[t]


@pytest.mark.parametrize("n", [1, 5, 16, 64, 100, 144, 257])
@pytest.mark.parametrize("ops", ["NN", "NC", "CN", "NT", "TN", "CC"])
def test_zgemm_matches_numpy(rng, n, ops):
    A, B, C0 = _rand(rng, n), _rand(rng, n), _rand(rng, n)
    alpha, beta = 0.7 - 0.2j, -1.0 + 0.5j
    got = capi.zgemm(ops[0], ops[1], A, B, alpha, beta, C0)
    ref = alpha * (_op(ops[0], A) @ _op(ops[1], B)) + beta * C0
    # FP64 dot products of length n: error <= n * eps * sum|a||b|
    bound = 4 * n * np.finfo(float).eps * (np.abs(_op(ops[0], A)) @ np.abs(_op(ops[1], B)) + np.abs(C0))
    assert np.all(np.abs(got - ref) <= bound + 1e-300)


# sizes cover one to eleven 256-row chunks of the inverse's tournament pivot
# selection, with full and partial last chunks, and partial outer blocks
@pytest.mark.parametrize("n", [1, 2, 7, 8, 9, 63, 64, 65, 100, 144, 256, 300, 512, 513, 1024, 1025,
                               1600, 2048, 2100, 2304, 2561, 2800])
def test_invert_matches_lapack(rng, n):
    A = _rand(rng, n) + 2.0 * np.sqrt(n) * np.eye(n)  # well conditioned
    X, info = capi.invert_matrix(A)
    assert info == 0
    ref = np.linalg.inv(A)
    rel = np.abs(X - ref).max() / np.abs(ref).max()
    assert rel < 1e-12, rel
    assert np.abs(A @ X - np.eye(n)).max() < 1e-11


@pytest.mark.parametrize("n", [96, 1500, 2500])
def test_invert_needs_pivoting(rng, n):
    # zero diagonal: elimination without row pivoting would divide by zero
    P = np.roll(np.eye(n), 7, axis=1)
    A = P * (1.0 + 0.3j) + 1e-3 * _rand(rng, n)
    np.fill_diagonal(A, 0.0)
    X, info = capi.invert_matrix(A)
    assert info == 0
    assert np.abs(A @ X - np.eye(n)).max() < 1e-10


@pytest.mark.parametrize("n", [64, 1024, 2048])
def test_invert_hadamard_ties(n):
    # every column offers many pivots of exactly equal magnitude; inv(c H) = H / (c n)
    H = np.ones((1, 1))
    while H.shape[0] < n:
        H = np.block([[H, H], [H, -H]])
    c = 1.0 + 1.0j
    X, info = capi.invert_matrix(c * H)
    assert info == 0
    # entries are 1/(|c| n); LAPACK's getrf/getri reproduces them exactly here
    assert np.abs(X - H.T / (c * n)).max() * n < 1e-10


@pytest.mark.parametrize("n", [70, 1100, 2400])
def test_invert_reports_singular(rng, n):
    A = _rand(rng, n)
    A[:, 33] = 0.0  # exactly singular: some pivot column is all zeros
    X, info = capi.invert_matrix(A)
    assert info > 0


@pytest.mark.parametrize("complex64,m3", [(True, 1), (True, 0), (True, 2), (False, 1), (False, 0)])
def test_bench_gemm_instrumentation_runs(complex64, m3):
    # negf_bench_gemm times the product kernel alone (counter runs); it must run
    # and report a positive launch time in every form
    ms = capi.bench_gemm(256, 3, complex64=complex64, m3=m3, reps=2)
    assert ms > 0.0


def test_bf16_split_form_is_complex64_only():
    with pytest.raises(capi.NegfError):
        capi.gemm_batch(np.zeros((1, 4, 4), complex), np.zeros((1, 4, 4), complex), complex64=False, m3=2)


# The complex64 product on the bf16 matrix cores (every FP32 operand split three
# ways, four real products): against numpy complex128 with the componentwise bound
# of the 4-product form, |Re err| <= 4 n eps32 (|Re A||Re B| + |Im A||Im B|) and
# |Im err| <= 4 n eps32 (|Re A||Im B| + |Im A||Re B|): its split products and FP32
# accumulation are at the FP32 rounding level.  Sizes: partial and whole 64x64
# tiles and 32-deep K slabs, up to the benchmark's block.
@pytest.mark.parametrize("n,batch", [(1, 2), (5, 2), (31, 3), (64, 2), (100, 2), (257, 2), (1024, 2), (2304, 3)])
def test_gemm_bf16_split_matches_numpy(rng, n, batch):
    print(f"bf16 split gemm n={n} batch={batch}", flush=True)
    A = (rng.standard_normal((batch, n, n)) + 1j * rng.standard_normal((batch, n, n))).astype(np.complex64)
    B = (rng.standard_normal((batch, n, n)) + 1j * rng.standard_normal((batch, n, n))).astype(np.complex64)
    A[0, :, 0] *= 1e-3  # a column of small entries: componentwise, not normwise
    got = capi.gemm_batch(A, B, complex64=True, m3=2)
    eps = np.finfo(np.float32).eps
    for b in range(batch):
        a, bb = A[b].astype(np.complex128), B[b].astype(np.complex128)
        ref = a @ bb
        ar, ai, br, bi = np.abs(a.real), np.abs(a.imag), np.abs(bb.real), np.abs(bb.imag)
        mre = ar @ br + ai @ bi
        mim = ar @ bi + ai @ br
        ere = np.abs(got[b].real - ref.real)
        eim = np.abs(got[b].imag - ref.imag)
        print(f"  b={b}: max err/bound re {(ere / (n * eps * mre)).max():.3f} im {(eim / (n * eps * mim)).max():.3f}",
              flush=True)
        assert np.all(ere <= 4 * n * eps * mre), (b, float((ere / mre).max() / eps))
        assert np.all(eim <= 4 * n * eps * mim), (b, float((eim / mim).max() / eps))


# The sweep's product kernel in each of its forms at the sweep's block sizes and
# batches: FP64 3M (RGF chains, trailing updates of the inverse, the mixed form's
# decimation), complex64 3M (the mixed form's first-column chain) and complex64
# 4-product, against numpy complex128.  Bounds: normwise for the 3M form, whose
# error is O(eps (|Re A|+|Im A|)(|Re B|+|Im B|)) per term, i.e. c * n * eps * |A||B|
# with |.| the elementwise sums |Re|+|Im|; componentwise for the 4-product form.
@pytest.mark.parametrize("n,batch", [(1024, 2), (1024, 8), (2304, 2), (2304, 8)])
@pytest.mark.parametrize("form", ["fp64_3m", "c64_3m", "c64_4p"])
def test_gemm_forms_at_sweep_sizes(rng, n, batch, form):
    print(f"gemm form {form} n={n} batch={batch}", flush=True)
    c64 = form.startswith("c64")
    m3 = form.endswith("3m")
    dt = np.complex64 if c64 else np.complex128
    A = (rng.standard_normal((batch, n, n)) + 1j * rng.standard_normal((batch, n, n))).astype(dt)
    B = (rng.standard_normal((batch, n, n)) + 1j * rng.standard_normal((batch, n, n))).astype(dt)
    got = capi.gemm_batch(A, B, complex64=c64, m3=m3)
    eps = np.finfo(np.float32 if c64 else np.float64).eps
    for b in (0, batch - 1):  # first and last matrix of the batch
        a, bb = A[b].astype(np.complex128), B[b].astype(np.complex128)
        ref = a @ bb
        mag = (np.abs(a.real) + np.abs(a.imag)) @ (np.abs(bb.real) + np.abs(bb.imag))
        err = np.abs(got[b] - ref)
        assert np.all(err <= 4 * n * eps * mag), (form, b, float((err / mag).max() / eps))


@pytest.mark.parametrize("n,batch", [(1024, 2), (2304, 2)])
@pytest.mark.parametrize("m3", [1, 0], ids=["3m", "4p"])
def test_invert_forms_at_sweep_sizes(rng, n, batch, m3):
    # the FP64 Gauss-Jordan inverse with 3M (RGF, mixed-form decimation) and 4-product
    # (plain-form decimation) trailing updates, against LAPACK
    print(f"invert m3={m3} n={n} batch={batch}", flush=True)
    G = np.stack([_rand(rng, n) + 2.0 * np.sqrt(n) * np.eye(n) for _ in range(batch)])
    X, info = capi.invert_batch(G, complex64=False, m3=bool(m3))
    assert np.all(info == 0)
    for b in range(batch):
        ref = np.linalg.inv(G[b])
        rel = np.abs(X[b] - ref).max() / np.abs(ref).max()
        assert rel < 1e-12, (b, rel)
        assert np.abs(G[b] @ X[b] - np.eye(n)).max() < 1e-11


def _block_shuffled(rng, n, kb, dt):
    # diagonally weighted (eigenvalues of R / sqrt(n) + 4 I lie in 4 +- 1.5), then the
    # rows of every kb-row diagonal block shuffled among themselves: the diagonal blocks
    # stay well conditioned, and the elimination needs row interchanges inside them
    G = _rand(rng, n) / np.sqrt(n) + 4.0 * np.eye(n)
    perm = np.concatenate([k0 + rng.permutation(min(kb, n - k0)) for k0 in range(0, n, kb)])
    return G[perm].astype(dt).astype(np.complex128)


@pytest.mark.parametrize("n", [1, 7, 64, 100, 128, 129, 300, 1024, 2304])
@pytest.mark.parametrize("c64,m3,group", [(True, 1, 1), (True, 2, 1), (False, 1, 1), (True, 1, 2), (True, 1, 4),
                                          (True, 2, 4)],
                         ids=["c64", "c64_bf16x3", "c128", "c64_group2", "c64_group4", "c64_bf16x3_group4"])
def test_invert_local_matches_lapack(rng, n, c64, m3, group):
    """negf_invert_batch_local (block Gauss-Jordan pivoting inside 128-column
    diagonal blocks in complex64, 64-column in complex128; the complex64 block-LU
    form's inverse) against LAPACK: one partial block up to the benchmark's 2304
    (18 and 36 blocks).  complex128: 1e-12 relative like the partially pivoted
    inverse; complex64: within 16x LAPACK's own complex64 inverse (cgetrf/cgetri
    through numpy) on the same matrix, or 16 eps; the trailing updates in the 3M
    form (m3 = 1) or on the bf16 matrix cores (m3 = 2); with group > 1 the blocks of
    each group of 128 group columns update only the group's columns and one
    rank-(128 group) product updates the rest (groups that are partial or the whole
    matrix included)."""
    kb = 128 if c64 else 64
    dt = np.complex64 if c64 else np.complex128
    batch = 2
    print(f"local inverse c64={c64} m3={m3} group={group} n={n} batch={batch}", flush=True)
    G = np.stack([_block_shuffled(rng, n, kb, dt) for _ in range(batch)])
    X, info = capi.invert_batch(G, complex64=c64, m3=m3, local=True, group=group)
    assert np.all(info == 0), info
    for b in range(batch):
        ref = np.linalg.inv(G[b])
        rel = np.abs(X[b] - ref).max() / np.abs(ref).max()
        if c64:
            lap = np.linalg.inv(G[b].astype(np.complex64)).astype(np.complex128)
            rel_lap = np.abs(lap - ref).max() / np.abs(ref).max()
            eps = np.finfo(np.float32).eps
            print(f"  n={n} b={b}: rel err {rel:.2e}, LAPACK complex64 {rel_lap:.2e}", flush=True)
            assert rel <= max(16 * rel_lap, 16 * eps), (b, rel, rel_lap)
        else:
            assert rel < 1e-12, (b, rel)
            assert np.abs(G[b] @ X[b] - np.eye(n)).max() < 1e-11


@pytest.mark.parametrize("c64,group", [(True, 1), (False, 1), (True, 2)], ids=["c64", "c128", "c64_group2"])
def test_invert_local_reports_zero_pivot_in_a_diagonal_block(rng, c64, group):
    # [[0, I], [I, 0]] is its own inverse, but its first diagonal block is zero: the
    # block Gauss-Jordan, which pivots only inside diagonal blocks, meets a zero pivot
    # in column 1 (info 1).  The second matrix of the batch is regular: its inverse and
    # info are unaffected.
    kb = 128 if c64 else 64
    n = 2 * kb
    swap = np.zeros((n, n), complex)
    swap[:kb, kb:] = np.eye(kb)
    swap[kb:, :kb] = np.eye(kb)
    good = _block_shuffled(rng, n, kb, np.complex64 if c64 else np.complex128)
    X, info = capi.invert_batch(np.stack([swap, good]), complex64=c64, m3=True, local=True, group=group)
    assert info.tolist() == [1, 0], info
    ref = np.linalg.inv(good)
    rel = np.abs(X[1] - ref).max() / np.abs(ref).max()
    assert rel < (1e-5 if c64 else 1e-12), rel
