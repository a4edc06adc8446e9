"""Problem descriptions: structured hex8 meshes (YaspGrid numbering), Dirichlet flags, traction faces
and point loads — the inputs that Ikarus assembles from makeBasis / DirichletValues / BoundaryPatch
(tests/src/testcantileverbeam.hh:102-151).

Numbering: vertex (i, j, k) -> i + (nx+1)(j + (ny+1)k), element likewise with nx, ny; element corner a
of the reference cube is (a&1, a>>1&1, a>>2&1) (Dune numbering); DOF = dim*vertex + component
(FlatInterleaved, ikarus/finiteelements/fehelper.hh:146-152).

Seeded interior jitter: SplitMix64 stream, one uniform per coordinate of each interior vertex in vertex
order, x_k += (2U - 1) * jitter * h_k. Boundary vertices stay on their planes.
"""
from dataclasses import dataclass, field

import numpy as np

GOLDEN_GAMMA = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed, count, offset=0):
    """The (offset+1)-th .. (offset+count)-th outputs of SplitMix64(seed) as uint64."""
    with np.errstate(over="ignore"):
        k = np.arange(offset + 1, offset + count + 1, dtype=np.uint64)
        z = np.uint64(seed) + k * GOLDEN_GAMMA
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        return z ^ (z >> np.uint64(31))


def uniforms(seed, count, offset=0):
    return (splitmix64(seed, count, offset) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def lame(E, nu):
    """toLamesFirstParameterAndShearModulus (ikarus/finiteelements/physicshelper.hh:276-281)."""
    return E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu)), E / (2.0 * (1.0 + nu))


def structured_mesh(nel, L, jitter=0.0, seed=0):
    nel = [int(n) for n in nel]
    dim = len(nel)
    nv = [n + 1 for n in nel]
    idx = np.indices(nv[::-1]).reshape(dim, -1)[::-1]  # idx[0] fastest
    h = np.array([L[k] / nel[k] for k in range(dim)])
    X = (idx.T * h).astype(np.float64)
    if jitter > 0.0:
        interior = np.ones(X.shape[0], dtype=bool)
        for k in range(dim):
            interior &= (idx[k] > 0) & (idx[k] < nel[k])
        ni = int(interior.sum())
        u = uniforms(seed, ni * dim).reshape(ni, dim)
        X[interior] += (2.0 * u - 1.0) * jitter * h
    ne = int(np.prod(nel))
    eidx = np.indices(nel[::-1]).reshape(dim, -1)[::-1]
    conn = np.zeros((ne, 1 << dim), dtype=np.int32)
    for a in range(1 << dim):
        off = [(a >> k) & 1 for k in range(dim)]
        v = np.zeros(ne, dtype=np.int64)
        stride = 1
        for k in range(dim):
            v += (eidx[k] + off[k]) * stride
            stride *= nv[k]
        conn[:, a] = v
    return X, conn


def face_elements(nel, axis, side):
    """Elements owning the boundary face normal to `axis` at the low (side=0) or high (side=1) end."""
    dim = len(nel)
    eidx = np.indices([int(n) for n in nel][::-1]).reshape(dim, -1)[::-1]
    target = 0 if side == 0 else nel[axis] - 1
    return np.nonzero(eidx[axis] == target)[0].astype(np.int32)


def clamp_flags(X, axis, value, comps=None, tol=1e-8):
    """Fix DOFs of vertices on the plane X[axis] == value (fixBoundaryDOFs, dirichletvalues.hh:111-131)."""
    dim = X.shape[1]
    on = np.abs(X[:, axis] - value) < tol
    flags = np.zeros(X.shape[0] * dim, dtype=np.uint8)
    for k in (range(dim) if comps is None else comps):
        flags[np.nonzero(on)[0] * dim + k] = 1
    return flags


def vertex_at(X, p, tol=1e-8):
    hit = np.nonzero(np.all(np.abs(X - np.asarray(p, dtype=np.float64)) < tol, axis=1))[0]
    if len(hit) != 1:
        raise ValueError(f"no unique vertex at {p}")
    return int(hit[0])


@dataclass
class Problem:
    X: np.ndarray
    conn: np.ndarray
    material: int              # 0 Neo-Hooke, 1 SVK
    lam: float
    mu: float
    eas: int = 0
    fixed: np.ndarray = None
    face_elem: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))
    face_axis: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int8))
    face_side: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int8))
    traction: tuple = (0.0, 0.0, 0.0)
    point_dof: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int64))
    point_coef: np.ndarray = field(default_factory=lambda: np.zeros(0))
    name: str = ""
    meta: dict = field(default_factory=dict)

    @property
    def dim(self):
        return self.X.shape[1]

    @property
    def ndof(self):
        return self.X.shape[0] * self.dim

    @property
    def ne(self):
        return self.conn.shape[0]

    def set_faces(self, elems, axis, side):
        self.face_elem = np.ascontiguousarray(elems, dtype=np.int32)
        self.face_axis = np.full(len(elems), axis, dtype=np.int8)
        self.face_side = np.full(len(elems), side, dtype=np.int8)


def load_params(seed, lo=10.0, hi=40.0):
    """Traction magnitude m ~ U[lo, hi] and direction uniform on the sphere, from the seed's stream."""
    u = uniforms(np.uint64(seed) ^ np.uint64(0x5EED5EED5EED5EED), 3)
    m = lo + (hi - lo) * u[0]
    z = 2.0 * u[1] - 1.0
    phi = 2.0 * np.pi * u[2]
    r = np.sqrt(max(0.0, 1.0 - z * z))
    return m, np.array([r * np.cos(phi), r * np.sin(phi), z])


def hex_block(n=(64, 64, 64), L=(1.0, 1.0, 1.0), material=0, E=1000.0, nu=0.3, jitter=0.05, seed=0,
              traction=None, mag_range=(10.0, 40.0), eas=0):
    """Unit block, bottom face z=0 clamped, traction on the top face z=L: the 3D hex8 benchmark problem."""
    X, conn = structured_mesh(n, L, jitter, seed)
    lam, mu = lame(E, nu)
    p = Problem(X, conn, material, lam, mu, eas=eas, name=f"hex{n[0]}x{n[1]}x{n[2]}")
    p.fixed = clamp_flags(X, 2, 0.0)
    p.set_faces(face_elements(n, 2, 1), 2, 1)
    if traction is None:
        m, d = load_params(seed, *mag_range)
        traction = m * d
    p.traction = tuple(float(t) for t in traction)
    p.meta = dict(seed=seed, jitter=jitter, E=E, nu=nu)
    return p


def bending_block(n=(96, 96, 48), L=(2.0, 2.0, 1.0), material=0, E=1000.0, nu=0.3, jitter=0.05, seed=0, eas=21,
                  mag_range=(2.0, 8.0)):
    """x=0 clamped, bending traction (0, 0, -m) on the x=L face."""
    X, conn = structured_mesh(n, L, jitter, seed)
    lam, mu = lame(E, nu)
    p = Problem(X, conn, material, lam, mu, eas=eas, name=f"eas{n[0]}x{n[1]}x{n[2]}")
    p.fixed = clamp_flags(X, 0, 0.0)
    p.set_faces(face_elements(n, 0, 1), 0, 1)
    m, _ = load_params(seed, *mag_range)
    p.traction = (0.0, 0.0, -float(m))
    p.meta = dict(seed=seed, jitter=jitter, E=E, nu=nu)
    return p


def cantilever2d(n=(64, 16), L=(4.0, 1.0), material=0, E=1000.0, nu=0.3, seed=0, mag_range=(0.5, 2.0)):
    """Plane-strain Q1 cantilever: x=0 edge clamped, uniform traction (0, -m) on the x=L edge with
    m ~ U[mag_range] from the seed's stream (64x16 elements over 4x1: the 2D Neo-Hooke benchmark problem)."""
    X, conn = structured_mesh(n, L)
    lam, mu = lame(E, nu)
    p = Problem(X, conn, material, lam, mu, name=f"quad{n[0]}x{n[1]}")
    p.fixed = clamp_flags(X, 0, 0.0)
    p.set_faces(face_elements(n, 0, 1), 0, 1)
    m, _ = load_params(seed, *mag_range)
    p.traction = (0.0, -float(m), 0.0)
    p.meta = dict(seed=seed, E=E, nu=nu)
    return p


def cantilever3d(material=0, E=100.0, nu=0.3, eas=21):
    """The reference's 3D cantilever: YaspGrid 10x1x1 over 10x2x2, x=0 clamped, unit point loads in y at
    (10,2,2) and (10,2,0) (tests/src/testcantileverbeam.hh:40-80,102-140)."""
    X, conn = structured_mesh((10, 1, 1), (10.0, 2.0, 2.0))
    lam, mu = lame(E, nu)
    p = Problem(X, conn, material, lam, mu, eas=eas, name="cantilever3d")
    p.fixed = clamp_flags(X, 0, 0.0)
    dofs = [vertex_at(X, (10, 2, 2)) * 3 + 1, vertex_at(X, (10, 2, 0)) * 3 + 1]
    p.point_dof = np.array(dofs, dtype=np.int64)
    p.point_coef = np.ones(2)
    return p
