"""Parity of the HIP path (through the C ABI) with the FP64 CPU oracle on the same seeded inputs.

Bars: sparsity pattern bit-exact; FP64 element/assembled values within 1e-10 relative (the summation
orders differ from the oracle's); converged displacements within 1e-8 relative.
"""
import numpy as np
import pytest

import ikarus_amd as ia
import pyoracle as po
from ikarus_amd import _capi as capi
from ikarus_amd import problems as pb

pytestmark = pytest.mark.gpu


def _oracle_for(p, neas=None):
    o = po.Oracle(p.X, p.conn, p.material, p.lam, p.mu, p.eas if neas is None else neas, 2)
    o.set_fixed(p.fixed)
    if len(p.face_elem):
        o.set_traction(p.face_elem, p.face_axis, p.face_side, p.traction)
    for d, c in zip(p.point_dof, p.point_coef):
        o.add_point_load(int(d), float(c))
    return o


def _relerr(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300)


@pytest.mark.parametrize("material", [capi.NEOHOOKE, capi.SVK])
def test_element_matrix_and_vector(material):
    p = pb.hex_block((3, 3, 3), material=material, jitter=0.08, seed=11)
    print(f"case: element parity material={material}", flush=True)
    dev = ia.Device(p)
    o = _oracle_for(p)
    u = 0.03 * np.random.default_rng(1).standard_normal(p.ndof)
    dev.set_requirement(u, 0.0)
    for e in (0, 13, 26):
        K, R = dev.element(e)
        Ko, Ro = o.element(e, u, 0.0)
        assert _relerr(K, Ko) < 1e-12, e
        assert _relerr(R, Ro) < 1e-12, e


@pytest.mark.parametrize("dbc", [capi.DBC_RAW, capi.DBC_FULL, capi.DBC_REDUCED])
@pytest.mark.parametrize("material", [capi.NEOHOOKE, capi.SVK])
def test_assembled_matrix_and_vector(dbc, material):
    p = pb.hex_block((6, 5, 4), material=material, jitter=0.05, seed=5)
    print(f"case: assembly parity dbc={dbc} material={material}", flush=True)
    dev = ia.Device(p)
    o = _oracle_for(p)
    u = 0.02 * np.random.default_rng(2).standard_normal(p.ndof)
    lam = 0.6
    dev.set_requirement(u, lam)
    dev.matrix(dbc)
    outer, inner, vals = dev.csc(dbc)
    oo, io = o.pattern(dbc == capi.DBC_REDUCED)
    assert np.array_equal(outer, oo) and np.array_equal(inner, io)
    Ko = o.matrix(u, lam, dbc)
    assert _relerr(vals, Ko) < 1e-10
    np.testing.assert_allclose(vals, Ko, rtol=1e-9, atol=1e-12 * np.abs(Ko).max())
    R = dev.vector(dbc)
    Ro = o.vector(u, lam, dbc)
    assert R.shape == Ro.shape
    assert _relerr(R, Ro) < 1e-10
    if dbc == capi.DBC_FULL:
        f = p.fixed.astype(bool)
        assert np.all(R[f] == 0.0)


def test_fp32_tangent_close_to_fp64():
    p = pb.hex_block((5, 5, 5), jitter=0.05, seed=9)
    print("case: fp32 tangent", flush=True)
    dev = ia.Device(p)
    u = 0.02 * np.random.default_rng(4).standard_normal(p.ndof) * (1 - p.fixed)
    dev.set_requirement(u, 1.0)
    dev.matrix(capi.DBC_FULL, capi.FP64)
    _, _, k64 = dev.csc()
    dev.matrix(capi.DBC_FULL, capi.FP32)
    _, _, k32 = dev.csc()
    assert _relerr(k32, k64) < 2e-6


def test_pattern_and_values_at_16cubed():
    p = pb.hex_block((16, 16, 16), jitter=0.05, seed=1)
    print("case: 16^3 parity", flush=True)
    dev = ia.Device(p)
    o = _oracle_for(p)
    u = 0.01 * np.random.default_rng(6).standard_normal(p.ndof)
    dev.set_requirement(u, 0.3)
    dev.matrix(capi.DBC_FULL)
    outer, inner, vals = dev.csc()
    oo, io = o.pattern()
    assert np.array_equal(outer, oo) and np.array_equal(inner, io)
    assert _relerr(vals, o.matrix(u, 0.3, po.FULL)) < 1e-10
    assert _relerr(dev.vector(), o.vector(u, 0.3, po.FULL)) < 1e-10


@pytest.mark.parametrize("kind", [capi.PCG_JACOBI_FP64, capi.PCG_BJACOBI_FP32_IR])
def test_linear_solve_matches_direct(kind):
    p = pb.hex_block((5, 4, 6), jitter=0.05, seed=3)
    print(f"case: linear solve kind={kind}", flush=True)
    dev = ia.Device(p)
    o = _oracle_for(p)
    u = 0.01 * np.random.default_rng(7).standard_normal(p.ndof) * (1 - p.fixed)
    dev.set_requirement(u, 0.5)
    dev.matrix(capi.DBC_FULL, capi.FP64 if kind == capi.PCG_JACOBI_FP64 else capi.FP32)
    b = np.random.default_rng(8).standard_normal(p.ndof) * (1 - p.fixed)
    x, info = dev.solve(b, ia.solver_opts(kind=kind, rtol=1e-12, inner_rtol=1e-3))
    xd = o.direct_solve(u, 0.5, b)
    assert info.converged == 1
    assert _relerr(x, xd) < (1e-9 if kind == capi.PCG_JACOBI_FP64 else 1e-5)
    if kind == capi.PCG_JACOBI_FP64:
        xc, it, _ = o.cg(u, 0.5, b, tol=1e-12)
        assert abs(info.iterations - it) <= 2  # Eigen ConjugateGradient iteration count
        assert _relerr(x, xc) < 1e-9


@pytest.mark.parametrize("material", [capi.NEOHOOKE, capi.SVK])
def test_load_control_matches_oracle(material):
    p = pb.hex_block((4, 4, 4), material=material, jitter=0.05, seed=21)
    print(f"case: load control material={material}", flush=True)
    dev = ia.Device(p)
    o = _oracle_for(p)
    r = o.load_control(3, 0.0, 1.0, tol=1e-8)
    dev.set_requirement(np.zeros(p.ndof), 0.0)
    cfg = capi.LCConfig()
    cfg.load_steps, cfg.tbegin, cfg.tend, cfg.predictor = 3, 0.0, 1.0, capi.PRED_CONSTANT
    info, iters, res = dev.load_control(cfg, ia.nr_settings(tol=1e-8, linear=ia.solver_opts(rtol=1e-13)))
    ug, lam = dev.get_requirement()
    assert info.success and abs(lam - 1.0) < 1e-14
    assert list(iters) == list(r["iterations"])
    assert np.all(res[1:] <= 1e-8)
    assert _relerr(ug, r["u"]) < 1e-8


def test_mixed_precision_newton_reaches_fp64_tolerance():
    p = pb.hex_block((6, 6, 6), jitter=0.05, seed=4)
    print("case: mixed newton", flush=True)
    o = _oracle_for(p)
    r = o.load_control(2, 0.0, 1.0, tol=1e-8)
    dev = ia.Device(p)
    dev.set_requirement(np.zeros(p.ndof), 0.0)
    cfg = capi.LCConfig()
    cfg.load_steps, cfg.tbegin, cfg.tend, cfg.predictor = 2, 0.0, 1.0, capi.PRED_LINEAR
    s = ia.nr_settings(tol=1e-8, precision=capi.FP32,
                       linear=ia.solver_opts(kind=capi.PCG_BJACOBI_FP32_IR, rtol=1e-10, inner_rtol=1e-3))
    info, iters, res = dev.load_control(cfg, s)
    ug, _ = dev.get_requirement()
    assert info.success and res[-1] <= 1e-8
    assert _relerr(ug, r["u"]) < 1e-8


def test_invalid_state_reports_element():
    """A non-finite displacement makes det C fail the positivity check: the call returns IK_ENEGDET and
    names the lowest element touching the node (the reference aborts, materialhelpers.hh:120-126)."""
    p = pb.hex_block((3, 3, 3), jitter=0.0, seed=0)
    print("case: invalid state", flush=True)
    dev = ia.Device(p)
    u = np.zeros(p.ndof)
    node = 21  # vertex (1,1,1)
    u[3 * node] = np.nan
    dev.set_requirement(u, 0.0)
    with pytest.raises(capi.IkError) as ei:
        dev.vector()
    assert ei.value.code == capi.IK_ENEGDET
    first = int(np.nonzero(np.any(p.conn == node, axis=1))[0].min())
    assert ei.value.bad_element == first
    # the handle stays usable
    dev.set_requirement(np.zeros(p.ndof), 0.0)
    assert np.all(dev.vector() == 0.0)


@pytest.mark.parametrize("kind", [capi.PCG_JACOBI_FP64, capi.PCG_TWOLEVEL_FP32_IR])
def test_invalid_state_reported_from_newton(kind):
    """Inside ik_newton_solve the assemblies leave their error-flag reads to the loop's next synchronisation:
    a non-finite displacement still makes the call return IK_ENEGDET naming the lowest element touching the
    node, and the same handle then solves a load step normally."""
    p = pb.hex_block((3, 3, 3), jitter=0.0, seed=0, traction=(0.0, 0.0, 0.05))
    print(f"case: invalid state in newton kind={kind}", flush=True)
    prec = capi.FP64 if kind == capi.PCG_JACOBI_FP64 else capi.FP32
    s = ia.nr_settings(tol=1e-8, max_iter=10, precision=prec, linear=ia.solver_opts(kind=kind, rtol=1e-10))
    dev = ia.Device(p)
    u = np.zeros(p.ndof)
    node = 21  # vertex (1,1,1)
    u[3 * node] = np.nan
    dev.set_requirement(u, 0.0)
    with pytest.raises(capi.IkError) as ei:
        dev.newton(s)
    assert ei.value.code == capi.IK_ENEGDET
    assert ei.value.bad_element == int(np.nonzero(np.any(p.conn == node, axis=1))[0].min())
    dev.set_requirement(np.zeros(p.ndof), 1.0)
    info = dev.newton(s)
    assert info.success == 1 and info.iterations >= 1 and info.residual_norm <= 1e-8
    dev.close()


def test_singular_eas_reported_from_newton():
    """The EAS elimination's pivot guard, read at the Newton loop's next synchronisation: IK_ESINGULAR naming
    element 0 when every D vanishes (lambda = mu = 0)."""
    p = pb.hex_block((2, 2, 2), material=capi.SVK, jitter=0.0, seed=0, eas=21, traction=(0.0, 0.0, 1.0))
    print("case: singular EAS D in newton", flush=True)
    p.lam, p.mu = 0.0, 0.0
    dev = ia.Device(p)
    dev.set_requirement(np.zeros(p.ndof), 1.0)
    with pytest.raises(capi.IkError) as ei:
        dev.newton(ia.nr_settings(tol=1e-8, max_iter=3, linear=ia.solver_opts(rtol=1e-10, max_iter=20)))
    assert ei.value.code == capi.IK_ESINGULAR and ei.value.bad_element == 0
    dev.close()


@pytest.mark.parametrize("eas", [0, 21])
def test_chunked_assembly_is_bitwise_identical(monkeypatch, eas):
    """Element chunks whose gathers overlap the next chunk on the second stream (IKARUS_AMD_CHUNK_ELEMENTS)
    give bitwise the same K and R as one chunk: every K entry and R component is summed in a fixed order."""
    p = pb.hex_block((7, 6, 5), jitter=0.05, seed=9, eas=eas)
    print(f"case: chunked assembly eas={eas}", flush=True)
    u = 0.01 * np.random.default_rng(4).standard_normal(p.ndof) * (1 - p.fixed)
    out = []
    for chunk in (None, "64", "100"):
        if chunk is None:
            monkeypatch.delenv("IKARUS_AMD_CHUNK_ELEMENTS", raising=False)
        else:
            monkeypatch.setenv("IKARUS_AMD_CHUNK_ELEMENTS", chunk)
        dev = ia.Device(p)
        dev.set_requirement(u, 0.7)
        dev.matrix(capi.DBC_FULL)
        _, _, vals = dev.csc()
        R = dev.vector()
        Rf = dev.assemble()  # fused pass: per-chunk residual gathers on the second stream
        _, _, valsf = dev.csc()
        out.append((vals, R, Rf, valsf))
        dev.close()
    for res in out[1:]:
        for a, b in zip(res, out[0]):
            assert np.array_equal(a, b)


@pytest.mark.parametrize("dbc", [capi.DBC_FULL, capi.DBC_REDUCED])
@pytest.mark.parametrize("eas", [0, 21])
def test_fused_assembly_matches_separate_calls(dbc, eas):
    """ik_assemble (K_e and R_e from one element pass) gives the residual of ik_vector and the matrix of
    ik_matrix up to summation order, and both match the oracle."""
    p = pb.hex_block((6, 5, 4), jitter=0.05, seed=8, eas=eas)
    print(f"case: fused assembly dbc={dbc} eas={eas}", flush=True)
    u = 0.01 * np.random.default_rng(3).standard_normal(p.ndof) * (1 - p.fixed)
    lam = 0.8
    dev = ia.Device(p)
    dev.set_requirement(u, lam)
    R1 = dev.vector(dbc)
    dev.matrix(dbc)
    _, _, K1 = dev.csc(dbc)
    R2 = dev.assemble(dbc)
    _, _, K2 = dev.csc(dbc)
    assert _relerr(R2, R1) < 1e-13
    if eas:
        assert np.array_equal(K2, K1)  # the EAS tangent always comes from the fused pass
    else:
        assert _relerr(K2, K1) < 1e-14
    o = _oracle_for(p)
    assert _relerr(R2, o.vector(u, lam, dbc)) < 1e-10
    assert _relerr(K2, o.matrix(u, lam, dbc)) < 1e-10
    if eas == 0:
        # FP32 tangent (plain element only): two passes, bitwise the results of the separate calls
        R3 = dev.assemble(dbc, capi.FP32)
        _, _, K3 = dev.csc(dbc)
        dev.matrix(dbc, capi.FP32)
        _, _, K4 = dev.csc(dbc)
        assert np.array_equal(K3, K4)
        assert np.array_equal(R3, R1)


def _scrambled_block(n, seed, drop=0.25):
    """A jittered block with a quarter of its elements removed, element order shuffled and nodes relabelled
    at random (orphan nodes dropped): no lexicographic numbering left for the tiling to rely on."""
    p = pb.hex_block(n, jitter=0.05, seed=seed, traction=(1.5, -2.0, 3.0))
    rng = np.random.default_rng(seed)
    keep = np.sort(rng.permutation(p.ne)[: int(round(p.ne * (1 - drop)))])
    order = rng.permutation(len(keep))
    old_elems = keep[order]                        # new element i is old element old_elems[i]
    conn = p.conn[old_elems]
    used = np.unique(conn)
    relabel = np.full(p.X.shape[0], -1, np.int64)
    relabel[used] = rng.permutation(len(used))     # new node id of each used old node
    X = np.zeros((len(used), 3))
    X[relabel[used]] = p.X[used]
    fixed = np.zeros(3 * len(used), np.uint8)
    fixed.reshape(-1, 3)[relabel[used]] = p.fixed.reshape(-1, 3)[used]
    new_of_old = np.full(p.ne, -1, np.int64)
    new_of_old[old_elems] = np.arange(len(old_elems))
    q = pb.Problem(X, relabel[conn].astype(np.int32), p.material, p.lam, p.mu, name="scrambled")
    q.fixed = fixed
    faces = new_of_old[p.face_elem]
    q.set_faces(np.sort(faces[faces >= 0]), int(p.face_axis[0]), int(p.face_side[0]))
    q.traction = p.traction
    return q


def test_scrambled_mesh_assembly(monkeypatch):
    """Parity of K (RAW and FULL) and R on a mesh with shuffled elements, relabelled nodes and holes, and
    bitwise the same results for chunked assembly (tiles, direct writes and row completion away from any
    lexicographic numbering)."""
    p = _scrambled_block((9, 8, 7), seed=17)
    print(f"case: scrambled mesh, {p.ne} elements, {p.X.shape[0]} nodes", flush=True)
    o = _oracle_for(p)
    u = 0.01 * np.random.default_rng(5).standard_normal(p.ndof) * (1 - p.fixed)
    lam = 0.9
    out = []
    for chunk in (None, "64"):
        if chunk is None:
            monkeypatch.delenv("IKARUS_AMD_CHUNK_ELEMENTS", raising=False)
        else:
            monkeypatch.setenv("IKARUS_AMD_CHUNK_ELEMENTS", chunk)
        dev = ia.Device(p)
        dev.set_requirement(u, lam)
        res = []
        for dbc in (capi.DBC_RAW, capi.DBC_FULL):
            R = dev.assemble(dbc)
            outer, inner, vals = dev.csc(dbc)
            if chunk is None:
                oo, io = o.pattern()
                assert np.array_equal(outer, oo) and np.array_equal(inner, io)
                assert _relerr(vals, o.matrix(u, lam, dbc)) < 1e-10
                assert _relerr(R, o.vector(u, lam, dbc)) < 1e-10
            res += [vals, R]
        out.append(res)
        dev.close()
    for a, b in zip(out[1], out[0]):
        assert np.array_equal(a, b)


def test_pattern_build_independent_of_thread_count(monkeypatch):
    """The host pattern build splits nodes, elements and tiles over threads; a 48^3 scrambled mesh (several
    chunks in every phase) built with 1 and with 5 threads gives the same pattern and bitwise the same K
    and R."""
    p = _scrambled_block((48, 48, 48), seed=23, drop=0.1)
    print(f"case: pattern threads, {p.ne} elements", flush=True)
    u = 0.01 * np.random.default_rng(9).standard_normal(p.ndof) * (1 - p.fixed)
    out = []
    for threads in ("1", "5"):
        monkeypatch.setenv("IK_PATTERN_THREADS", threads)
        dev = ia.Device(p)
        dev.set_requirement(u, 0.7)
        R = dev.assemble(capi.DBC_FULL)
        out.append((dev.csc(capi.DBC_FULL), R, dev.assembly_layout()))
        dev.close()
    (c1, R1, l1), (c5, R5, l5) = out
    assert all(np.array_equal(a, b) for a, b in zip(c1, c5))
    assert np.array_equal(R1, R5) and l1 == l5


@pytest.mark.parametrize("material", [capi.NEOHOOKE, capi.SVK])
def test_tangent_is_derivative_of_residual(material):
    """The device's own consistency, independent of the oracle (checkfebyautodiff.hh:19-115,
    testnonlinearelasticity.hh:194-246): the RAW tangent equals the central difference of the RAW residual,
    is symmetric, and at u = 0 has the 6 rigid-body modes of the free body in its null space."""
    import scipy.sparse as sp
    p = pb.hex_block((3, 3, 2), material=material, jitter=0.08, seed=61, traction=(2.0, -1.0, 3.0))
    print(f"case: tangent consistency material={material}", flush=True)
    dev = ia.Device(p)
    n = p.ndof
    rng = np.random.default_rng(8)
    u = 0.02 * rng.standard_normal(n)
    lam = 0.6

    def tangent(uu):
        dev.set_requirement(uu, lam)
        dev.matrix(capi.DBC_RAW)
        outer, inner, vals = dev.csc(capi.DBC_RAW)
        return sp.csc_matrix((vals, inner, outer), shape=(n, n)).toarray()

    K = tangent(u)
    assert np.abs(K - K.T).max() <= 1e-13 * np.abs(K).max()
    h = 1e-6
    for j in rng.choice(n, 8, replace=False):
        e = np.zeros(n)
        e[j] = h
        dev.set_requirement(u + e, lam)
        Rp = dev.vector(capi.DBC_RAW)
        dev.set_requirement(u - e, lam)
        Rm = dev.vector(capi.DBC_RAW)
        assert np.abs((Rp - Rm) / (2 * h) - K[:, j]).max() <= 1e-6 * np.abs(K).max()
    K0 = tangent(np.zeros(n))
    X = p.X
    modes = []
    for k in range(3):
        t = np.zeros((X.shape[0], 3))
        t[:, k] = 1.0
        modes.append(t.ravel())
    for a, b in ((0, 1), (1, 2), (2, 0)):
        r = np.zeros((X.shape[0], 3))
        r[:, a], r[:, b] = -X[:, b], X[:, a]
        modes.append(r.ravel())
    for m in modes:
        assert np.linalg.norm(K0 @ m) <= 1e-12 * np.abs(K0).max() * np.linalg.norm(m)
    dev.close()


@pytest.mark.parametrize("precision", [capi.FP64, capi.FP32])
def test_block_view_matches_oracle(precision):
    """ik_matrix_bsr's device view (block CSR pattern, each block at values + 9 block_pos[b]) rebuilds the FULL
    matrix of the oracle, on the scrambled mesh (sliced-ELL slices of uneven width) and a structured block."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]

    def d2h(ptr, dtype, n):
        a = np.zeros(n, dtype=dtype)
        if n:
            assert hip.hipMemcpy(a.ctypes.data_as(C.c_void_p), ptr, a.nbytes, 2) == 0
        return a

    for p in (_scrambled_block((9, 8, 7), seed=23), pb.hex_block((6, 5, 4), seed=2)):
        print(f"case: block view, {p.ne} elements, precision {precision}", flush=True)
        o = _oracle_for(p)
        u = 0.01 * np.random.default_rng(8).standard_normal(p.ndof) * (1 - p.fixed)
        dev = ia.Device(p)
        dev.set_requirement(u, 0.7)
        dev.check(capi.lib().ik_matrix(dev.h, capi.DBC_FULL, precision))
        nbr, nnzb = C.c_int64(), C.c_int64()
        rp, ci, bp, vals = C.c_void_p(), C.c_void_p(), C.c_void_p(), C.c_void_p()
        prec = C.c_int()
        dev.check(capi.lib().ik_matrix_bsr(dev.h, C.byref(nbr), C.byref(nnzb), C.byref(rp), C.byref(ci), C.byref(bp),
                                           C.byref(vals), C.byref(prec)))
        assert nbr.value == p.X.shape[0] and prec.value == precision
        rowptr = d2h(rp, np.int32, nbr.value + 1)
        colidx = d2h(ci, np.int32, nnzb.value)
        pos = d2h(bp, np.int32, nnzb.value)
        assert rowptr[-1] == nnzb.value and len(np.unique(pos)) == nnzb.value and pos.min() >= 0
        dt = np.float64 if precision == capi.FP64 else np.float32
        V = d2h(vals, dt, 9 * (int(pos.max()) + 1)).reshape(-1, 3, 3)[pos].astype(np.float64)
        n = 3 * nbr.value
        K = np.zeros((n, n))
        for i in range(nbr.value):
            for b in range(rowptr[i], rowptr[i + 1]):
                j = colidx[b]
                K[3 * i:3 * i + 3, 3 * j:3 * j + 3] = V[b]
        Ko = o.matrix_dense(u, 0.7, po.FULL)
        tol = 1e-10 if precision == capi.FP64 else 1e-6
        assert np.abs(K - Ko).max() <= tol * np.abs(Ko).max()
        dev.close()
