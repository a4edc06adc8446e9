"""Host logic and the C-ABI boundary without a GPU: mesh generation parity with the oracle, problem
builders, and that libikarus_amd.so loads and exports every symbol of include/ikarus_amd.h."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import pyoracle as po
from ikarus_amd import _capi as capi
from ikarus_amd import problems as pb

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.parametrize("nel,L,jit,seed", [([3, 2, 4], [1.0, 2.0, 0.5], 0.05, 7), ([5, 3], [4.0, 1.0], 0.0, 0),
                                            ([4, 4, 4], [1.0, 1.0, 1.0], 0.05, 123456789)])
def test_mesh_matches_oracle_bitwise(nel, L, jit, seed):
    X, conn = pb.structured_mesh(nel, L, jit, seed)
    Xo, co = po.structured_mesh(nel, L, jit, seed)
    assert np.array_equal(conn, co)
    assert np.array_equal(X, Xo)


def test_splitmix_known_values():
    # SplitMix64 reference outputs for seed 0 (Steele, Lea, Flood 2014; as in java.util.SplittableRandom)
    v = pb.splitmix64(0, 3)
    assert [int(x) for x in v] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


def test_jitter_interior_only_and_bounded():
    X, _ = pb.structured_mesh([6, 5, 4], [1.0, 1.0, 1.0], 0.05, 3)
    X0, _ = pb.structured_mesh([6, 5, 4], [1.0, 1.0, 1.0], 0.0, 3)
    d = X - X0
    h = np.array([1 / 6, 1 / 5, 1 / 4])
    assert np.all(np.abs(d) <= 0.05 * h + 1e-15)
    for k in range(3):
        on = (np.abs(X0[:, k]) < 1e-12) | (np.abs(X0[:, k] - 1.0) < 1e-12)
        assert np.all(d[on] == 0.0)


def test_hex_block_problem():
    p = pb.hex_block((4, 4, 4), seed=2)
    assert p.fixed.sum() == 25 * 3
    assert len(p.face_elem) == 16 and np.all(p.face_axis == 2) and np.all(p.face_side == 1)
    m = np.linalg.norm(p.traction)
    assert 10.0 <= m <= 40.0
    c = pb.cantilever3d()
    assert c.fixed.sum() == 4 * 3 and len(c.point_dof) == 2


def test_library_exports_header_symbols():
    lib = capi.lib()
    header = open(os.path.join(ROOT, "include", "ikarus_amd.h")).read()
    names = set(re.findall(r"\b(ik_[a-z0-9_]+)\s*\(", header))
    assert names, "no declarations found"
    assert names == set(capi.EXPORTED), names ^ set(capi.EXPORTED)
    for n in sorted(names):
        assert hasattr(lib, n), n
    assert lib.ik_mlp_num_params() == 14 * 128 + 128 + 2 * (128 * 128 + 128) + 3 * 128 + 3


def test_capi_rejects_bad_input_without_gpu_compute():
    """Validation happens before any device work: an unsupported dimension, EAS modes on a plane-strain mesh,
    out-of-range connectivity and a 3D traction face on a 2D mesh are refused."""
    lib = capi.lib()
    X = np.zeros((4, 2))
    conn = np.array([[0, 1, 2, 3]], np.int32)

    def create(dim=2, eas=0, c=conn, faces=None):
        d = capi.ProblemDesc()
        d.dim = dim
        d.num_nodes, d.num_elements = 4, 1
        d.X, d.conn = capi.ptr(X), capi.ptr(c)
        d.eas_modes = eas
        keep = []
        if faces is not None:
            fe, fa, fs = (np.array([0], np.int32), np.array([faces], np.int8), np.array([1], np.int8))
            keep += [fe, fa, fs]
            d.num_faces, d.face_elem, d.face_axis, d.face_side = 1, capi.ptr(fe), capi.ptr(fa), capi.ptr(fs)
        h = C.c_void_p()
        rc = lib.ik_create(C.byref(h), C.byref(d))
        msg = lib.ik_last_error(h)
        lib.ik_destroy(h)
        return rc, msg

    rc, msg = create(dim=4)
    assert rc == capi.IK_ENOTIMPL and b"dim must be 3" in msg
    rc, msg = create(eas=4)
    assert rc == capi.IK_ENOTIMPL and b"EAS" in msg
    rc, msg = create(c=np.array([[0, 1, 2, 7]], np.int32))
    assert rc == capi.IK_EINVAL and b"out of range" in msg
    rc, msg = create(faces=2)
    assert rc == capi.IK_EINVAL and b"2D mesh" in msg


def test_pattern_build_under_sanitizers():
    """The threaded host pattern build, run on the CPU under ASan + UBSan and under TSan (tests/cpp/
    pattern_sanitize.cpp: HIP runtime calls replaced by host test doubles; 1, 3 and 8 threads byte-identical)."""
    import subprocess
    import __graft_entry__ as ge
    out = os.path.join(ROOT, "tests", "cpp", "_build")
    exes = [os.path.join(out, n) for n in ("pattern_asan", "pattern_tsan")]
    if not all(os.path.exists(e) for e in exes):
        exes = ge.build_sanitizer_harness()
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1", TSAN_OPTIONS="halt_on_error=1")
    for exe in exes:
        r = subprocess.run([exe], capture_output=True, text=True, env=env, timeout=300)
        assert r.returncode == 0 and r.stdout.strip().endswith("OK"), (exe, r.stdout[-2000:], r.stderr[-4000:])


def test_capi_rejects_degenerate_connectivity_without_gpu_compute():
    """A node in no element (no Dune grid has one; its block row would be empty) and an element that repeats a node
    are refused before any device work."""
    lib = capi.lib()
    X = np.zeros((9, 3))
    X[:8] = [[a & 1, (a >> 1) & 1, (a >> 2) & 1] for a in range(8)]

    def create(conn):
        d = capi.ProblemDesc()
        d.dim = 3
        d.num_nodes, d.num_elements = len(X), 1
        d.X, d.conn = capi.ptr(X), capi.ptr(conn)
        h = C.c_void_p()
        rc = lib.ik_create(C.byref(h), C.byref(d))
        msg = lib.ik_last_error(h)
        lib.ik_destroy(h)
        return rc, msg

    rc, msg = create(np.arange(8, dtype=np.int32))
    assert rc == capi.IK_EINVAL and b"node 8 belongs to no element" in msg
    rc, msg = create(np.array([0, 1, 2, 3, 4, 5, 6, 6], np.int32))
    assert rc == capi.IK_EINVAL and b"repeats a node" in msg
