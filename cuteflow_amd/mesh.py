"""Triangle meshes: the reference's text mesh format, the boundary table,
seeded synthetic generators and cell reorderings.

Text mesh layout follows lec_mail_tri (maillage.cuf:196-396); the boundary
table follows the `boundary_table_<meshfile>` file of maillage.cuf:55-79.
"""
from __future__ import annotations

import dataclasses

import numpy as np

from . import _lib


@dataclasses.dataclass
class Mesh:
    coords: np.ndarray            # (nnt, 3) x, y, z
    conn: np.ndarray              # (nelt, 3) 1-based node ids, counter-clockwise
    manning: np.ndarray           # (nelt,) per-cell Manning n (manning_nelt)
    ndinput: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros(0, np.int32))
    numero_ndinput: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros(0, np.int32))
    ndoutput: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros(0, np.int32))
    numero_ndoutput: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros(0, np.int32))
    ndwall: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros(0, np.int32))
    manning_nd: np.ndarray | None = None
    nombre_entrees: int = 1
    nombre_sorties: int = 1

    def __post_init__(self):
        self.coords = np.ascontiguousarray(self.coords, dtype=np.float64)
        self.conn = np.ascontiguousarray(self.conn, dtype=np.int32)
        self.manning = np.ascontiguousarray(self.manning, dtype=np.float64)
        for f in ("ndinput", "numero_ndinput", "ndoutput", "numero_ndoutput", "ndwall"):
            setattr(self, f, np.ascontiguousarray(getattr(self, f), dtype=np.int32))
        if len(self.numero_ndinput) != len(self.ndinput):
            self.numero_ndinput = np.ones(len(self.ndinput), np.int32)
        if len(self.numero_ndoutput) != len(self.ndoutput):
            self.numero_ndoutput = np.ones(len(self.ndoutput), np.int32)

    @property
    def nelt(self) -> int:
        return int(self.conn.shape[0])

    @property
    def nnt(self) -> int:
        return int(self.coords.shape[0])

    def centroids(self):
        """Cell centroids as construct_coo_table_elemwise (maillage.cuf:1271-1272)."""
        x = self.coords[:, 0][self.conn - 1]
        y = self.coords[:, 1][self.conn - 1]
        return ((x[:, 0] + x[:, 1]) + x[:, 2]) / 3.0, ((y[:, 0] + y[:, 1]) + y[:, 2]) / 3.0

    def boundary_table(self, device: bool = False) -> np.ndarray:
        """boundary(nelt,3,2) as a (2, 3, nelt) int32 array (the Fortran
        column-major layout), built by the library: cf_build_boundary on the
        host, or cf_build_boundary_device on the GPU (same table)."""
        lib = _lib.load()
        bnd = np.zeros((2, 3, self.nelt), np.int32)
        conn = np.ascontiguousarray(self.conn.reshape(-1))
        fn = lib.cf_build_boundary_device if device else lib.cf_build_boundary
        st = fn(self.nelt, self.nnt, _lib.iptr(conn),
                len(self.ndinput), _lib.iptr(self.ndinput), _lib.iptr(self.numero_ndinput),
                len(self.ndoutput), _lib.iptr(self.ndoutput), _lib.iptr(self.numero_ndoutput),
                len(self.ndwall), _lib.iptr(self.ndwall), _lib.iptr(bnd))
        _lib.check(st, "cf_build_boundary_device" if device else "cf_build_boundary")
        return bnd

    def inlet_lengths(self, bnd: np.ndarray, nentrees: int | None = None) -> np.ndarray:
        """long_entree (maillage.cuf:1423-1452), via cf_inlet_lengths."""
        lib = _lib.load()
        n = self.nombre_entrees if nentrees is None else nentrees
        out = np.zeros(_lib.CF_MAX_IO, np.float64)
        conn = np.ascontiguousarray(self.conn.reshape(-1))
        coords = np.ascontiguousarray(self.coords.reshape(-1))
        b = np.ascontiguousarray(bnd, dtype=np.int32)
        st = lib.cf_inlet_lengths(self.nelt, self.nnt, _lib.iptr(conn), _lib.dptr(coords),
                                  _lib.iptr(b), n, _lib.dptr(out))
        _lib.check(st, "cf_inlet_lengths")
        return out

    def permuted(self, perm: np.ndarray) -> "Mesh":
        """Mesh whose cell k is this mesh's cell perm[k]."""
        perm = np.asarray(perm)
        return dataclasses.replace(self, conn=self.conn[perm].copy(), manning=self.manning[perm].copy())


# ---------------------------------------------------------------------------
# text formats

def _numbers(lines, pos: int, count: int, conv=float):
    """List-directed read of `count` values starting at line `pos` (values may
    continue on the next lines); returns (values, next line)."""
    vals = []
    while len(vals) < count:
        if pos >= len(lines):
            raise ValueError("unexpected end of mesh file")
        vals += lines[pos].replace(",", " ").split()
        pos += 1
    return [conv(v.lower().replace("d", "e")) if conv is float else conv(v) for v in vals[:count]], pos


def read_text_mesh(path: str, nombre_entrees: int = 1, nombre_sorties: int = 1) -> Mesh:
    """Read a mesh in the text layout of lec_mail_tri (maillage.cuf:196-396),
    single-subdomain form."""
    with open(path, "r", encoding="latin-1") as fh:
        lines = fh.read().splitlines()
    pos = 1                                            # title
    (nnt,), pos = _numbers(lines, pos, 1, int)
    coords = np.zeros((nnt, 3))
    mnd = np.zeros(nnt)
    for i in range(nnt):
        v, pos = _numbers(lines, pos, 5)
        coords[i] = v[1:4]
        mnd[i] = v[4]
    pos += 1                                           # title
    (nelt,), pos = _numbers(lines, pos, 1, int)
    conn = np.zeros((nelt, 3), np.int32)
    mann = np.zeros(nelt)
    for i in range(nelt):
        v, pos = _numbers(lines, pos, 5)
        conn[i] = [int(v[1]), int(v[2]), int(v[3])]
        mann[i] = v[4]

    def node_block(pos, n_io):
        pos += 1                                       # title
        if n_io == 1:
            (nd,), pos = _numbers(lines, pos, 1, int)
        else:
            (nd, n_io), pos = _numbers(lines, pos, 2, int)
        ids, num = np.zeros(nd, np.int32), np.ones(nd, np.int32)
        for i in range(nd):
            if n_io == 1:
                (ids[i],), pos = _numbers(lines, pos, 1, int)
            else:
                (ids[i], num[i]), pos = _numbers(lines, pos, 2, int)
        return ids, num, n_io, pos

    ndi, numi, nombre_entrees, pos = node_block(pos, nombre_entrees)
    ndo, numo, nombre_sorties, pos = node_block(pos, nombre_sorties)
    pos += 1
    (ndw,), pos = _numbers(lines, pos, 1, int)
    ndwall = np.zeros(ndw, np.int32)
    for i in range(ndw):
        (ndwall[i],), pos = _numbers(lines, pos, 1, int)
    return Mesh(coords=coords, conn=conn, manning=mann, ndinput=ndi, numero_ndinput=numi,
                ndoutput=ndo, numero_ndoutput=numo, ndwall=ndwall, manning_nd=mnd,
                nombre_entrees=nombre_entrees, nombre_sorties=nombre_sorties)


def write_text_mesh(path: str, m: Mesh) -> None:
    """Write the layout read by read_text_mesh (full double precision)."""
    mnd = m.manning_nd if m.manning_nd is not None else np.zeros(m.nnt)
    with open(path, "w") as fh:
        fh.write("Table des coordonnees des noeuds\n")
        fh.write(f"{m.nnt}\n")
        for i in range(m.nnt):
            x, y, z = m.coords[i]
            fh.write(f"{i + 1} {float(x)!r} {float(y)!r} {float(z)!r} {float(mnd[i])!r}\n")
        fh.write("Table de connectivite\n")
        fh.write(f"{m.nelt}\n")
        for i in range(m.nelt):
            a, b, c = m.conn[i]
            fh.write(f"{i + 1} {a} {b} {c} {float(m.manning[i])!r}\n")
        for title, ids, num, n_io in (("Noeuds d'entree", m.ndinput, m.numero_ndinput, m.nombre_entrees),
                                      ("Noeuds de sortie", m.ndoutput, m.numero_ndoutput, m.nombre_sorties)):
            fh.write(title + "\n")
            if n_io == 1:
                fh.write(f"{len(ids)}\n")
                for v in ids:
                    fh.write(f"{v}\n")
            else:
                fh.write(f"{len(ids)} {n_io}\n")
                for v, k in zip(ids, num):
                    fh.write(f"{v} {k}\n")
        fh.write("Noeuds de murs\n")
        fh.write(f"{len(m.ndwall)}\n")
        for v in m.ndwall:
            fh.write(f"{v}\n")


def write_boundary_table(path: str, bnd: np.ndarray) -> None:
    """boundary_table file: one line per cell, b(i,1,1) b(i,2,1) b(i,3,1)
    b(i,1,2) b(i,2,2) b(i,3,2) (maillage.cuf:65-67)."""
    n = bnd.shape[2]
    rows = np.concatenate([bnd[0].T, bnd[1].T], axis=1)
    with open(path, "w") as fh:
        for r in rows:
            fh.write(" " + " ".join(f"{int(v):12d}" for v in r) + "\n")
    assert rows.shape == (n, 6)


def read_boundary_table(path: str, nelt: int | None = None) -> np.ndarray:
    """Inverse of write_boundary_table (maillage.cuf:71-75)."""
    rows = np.loadtxt(path, dtype=np.int64, ndmin=2)
    if nelt is not None:
        rows = rows[:nelt]
    bnd = np.zeros((2, 3, rows.shape[0]), np.int32)
    bnd[0] = rows[:, 0:3].T
    bnd[1] = rows[:, 3:6].T
    return bnd


def read_vtk_mesh(path: str) -> Mesh:
    """ASCII legacy-VTK UNSTRUCTURED_GRID of triangles with an optional cell
    scalar (the `mannings.vtk` files of the calibration workflow)."""
    with open(path, "r", encoding="latin-1") as fh:
        toks = fh.read().split()
    i = toks.index("POINTS")
    nnt = int(toks[i + 1])
    pts = np.array(toks[i + 3:i + 3 + 3 * nnt], dtype=np.float64).reshape(nnt, 3)
    i = toks.index("CELLS")
    nelt = int(toks[i + 1])
    cells = np.array(toks[i + 3:i + 3 + 4 * nelt], dtype=np.int64).reshape(nelt, 4)
    if not np.all(cells[:, 0] == 3):
        raise ValueError("only triangles are supported")
    mann = np.zeros(nelt)
    if "CELL_DATA" in toks:
        j = toks.index("LOOKUP_TABLE", toks.index("CELL_DATA"))
        mann = np.array(toks[j + 2:j + 2 + nelt], dtype=np.float64)
    return Mesh(coords=pts, conn=(cells[:, 1:] + 1).astype(np.int32), manning=mann)


# ---------------------------------------------------------------------------
# synthetic generators

def _grid_nodes(nx, ny, lx, ly, jitter, rng, x0=0.0, y0=0.0):
    xs = np.linspace(x0, x0 + lx, nx + 1)
    ys = np.linspace(y0, y0 + ly, ny + 1)
    X, Y = np.meshgrid(xs, ys, indexing="xy")   # (ny+1, nx+1)
    dx, dy = lx / nx, ly / ny
    if jitter > 0:
        jx = rng.uniform(-jitter, jitter, X.shape) * dx
        jy = rng.uniform(-jitter, jitter, Y.shape) * dy
        interior = np.zeros(X.shape, bool)
        interior[1:-1, 1:-1] = True
        X = X + np.where(interior, jx, 0.0)
        Y = Y + np.where(interior, jy, 0.0)
    return X.reshape(-1), Y.reshape(-1)


def _grid_triangles(nx, ny):
    """Two counter-clockwise triangles per quad, diagonal alternating in a
    checkerboard; node (i, j) has id j*(nx+1)+i+1."""
    i, j = np.meshgrid(np.arange(nx), np.arange(ny), indexing="xy")
    i, j = i.reshape(-1), j.reshape(-1)
    a = j * (nx + 1) + i + 1
    b = a + 1
    d = a + (nx + 1)
    c = d + 1
    even = ((i + j) % 2) == 0
    t1 = np.where(even[:, None], np.stack([a, b, c], 1), np.stack([a, b, d], 1))
    t2 = np.where(even[:, None], np.stack([a, c, d], 1), np.stack([b, c, d], 1))
    tri = np.empty((2 * len(a), 3), np.int64)
    tri[0::2] = t1
    tri[1::2] = t2
    return tri.astype(np.int32)


def _boundary_nodes(nx, ny):
    ids = np.arange((nx + 1) * (ny + 1)).reshape(ny + 1, nx + 1) + 1
    return ids, np.unique(np.concatenate([ids[0], ids[-1], ids[:, 0], ids[:, -1]])).astype(np.int32)


def dam_break_mesh(n: int = 100, size: float = 200.0, jitter: float = 0.2, seed: int = 0) -> Mesh:
    """Square [0,size]^2, n x n quads, interior nodes jittered by
    U(-jitter, jitter) cells, 2 triangles per quad, flat bed, n = 0,
    walls all around."""
    rng = np.random.default_rng(seed)
    x, y = _grid_nodes(n, n, size, size, jitter, rng)
    coords = np.stack([x, y, np.zeros_like(x)], 1)
    conn = _grid_triangles(n, n)
    _, walls = _boundary_nodes(n, n)
    return Mesh(coords=coords, conn=conn, manning=np.zeros(len(conn)), ndwall=walls,
                nombre_entrees=1, nombre_sorties=1)


def circular_dam_state(m: Mesh, center=(100.0, 100.0), radius: float = 30.0,
                       h_in: float = 10.0, h_out: float = 1.0) -> np.ndarray:
    """Initial state (nelt, 3): h_in inside the circle (by centroid), h_out
    outside, at rest."""
    xc, yc = m.centroids()
    inside = (xc - center[0]) ** 2 + (yc - center[1]) ** 2 <= radius * radius
    U = np.zeros((m.nelt, 3))
    U[:, 0] = np.where(inside, h_in, h_out)
    return U


def smooth_random_field(x, y, sigma, corr_len, seed, extent):
    """Seeded Gaussian-smoothed white noise of standard deviation `sigma` and
    correlation length `corr_len`, sampled bilinearly at (x, y)."""
    from scipy import ndimage

    x0, y0, x1, y1 = extent
    h = corr_len / 4.0
    nx = int(np.ceil((x1 - x0) / h)) + 9
    ny = int(np.ceil((y1 - y0) / h)) + 9
    rng = np.random.default_rng(seed)
    noise = rng.standard_normal((ny, nx))
    field = ndimage.gaussian_filter(noise, sigma=corr_len / h / 2.0, mode="wrap")
    field *= sigma / field.std()
    gx = (np.asarray(x) - x0) / h + 4.0
    gy = (np.asarray(y) - y0) / h + 4.0
    return ndimage.map_coordinates(field, [gy, gx], order=1, mode="nearest")


@dataclasses.dataclass
class River:
    mesh: Mesh
    state: np.ndarray       # initial (nelt, 3)
    channel: np.ndarray     # bool per cell
    h_out: float            # outlet water level
    q_peak: float
    t_ramp: float


def river_mesh(nx: int = 3200, ny: int = 320, length: float = 20000.0, width: float = 2000.0,
               channel_width: float = 200.0, depth: float = 3.0, slope: float = 1.0 / 2000.0,
               bar_amp: float = 0.5, bar_len: float = 300.0, field_sigma: float = 0.3,
               field_corr: float = 200.0, n_channel: float = 0.03, n_plain: float = 0.06,
               q_peak: float = 3000.0, t_ramp: float = 3600.0, jitter: float = 0.2,
               bank_width: float | None = None, seed: int = 1) -> River:
    """River reach with a floodplain: a `depth`-deep main channel along the
    centre line of a dry floodplain; bed = -slope*x + sinusoidal bars +
    seeded smooth random field; inflow on the channel section at x = 0,
    fixed water level over the whole section at x = length, walls elsewhere.

    The channel banks slope linearly over `bank_width` (default 8 cells
    across): a one-cell cliff makes the reference's WAF scheme (superbee
    limiter) blow up where shallow floodplain water flows over the bank."""
    rng = np.random.default_rng(seed)
    x, y = _grid_nodes(nx, ny, length, width, jitter, rng)
    yc0 = width / 2.0
    if bank_width is None:
        bank_width = 8.0 * width / ny
    z = -slope * x + bar_amp * np.sin(2.0 * np.pi * x / bar_len)
    z = z + smooth_random_field(x, y, field_sigma, field_corr, seed + 1000, (0.0, 0.0, length, width))
    frac = np.clip((channel_width / 2.0 + bank_width / 2.0 - np.abs(y - yc0)) / bank_width, 0.0, 1.0)
    z = z - depth * frac
    coords = np.stack([x, y, z], 1)
    conn = _grid_triangles(nx, ny)
    ids, bnodes = _boundary_nodes(nx, ny)
    inlet = ids[:, 0][np.abs(y[ids[:, 0] - 1] - yc0) <= channel_width / 2.0 + 1e-9].astype(np.int32)
    outlet = ids[:, -1].astype(np.int32)
    walls = np.setdiff1d(bnodes, np.concatenate([inlet, outlet])).astype(np.int32)
    xc = ((x[conn[:, 0] - 1] + x[conn[:, 1] - 1]) + x[conn[:, 2] - 1]) / 3.0
    ycc = ((y[conn[:, 0] - 1] + y[conn[:, 1] - 1]) + y[conn[:, 2] - 1]) / 3.0
    channel = np.abs(ycc - yc0) <= channel_width / 2.0
    mann = np.where(channel, n_channel, n_plain)
    m = Mesh(coords=coords, conn=conn, manning=mann, ndinput=inlet, ndoutput=outlet, ndwall=walls,
             nombre_entrees=1, nombre_sorties=1)
    # initial water surface 1.5 m below the mean floodplain level: the
    # channel holds water, the floodplain starts dry
    eta0 = -slope * xc - 1.5
    zc = (z[conn[:, 0] - 1] + z[conn[:, 1] - 1] + z[conn[:, 2] - 1]) / 3.0
    h0 = np.where(channel, np.maximum(eta0 - zc, 0.0), 0.0)
    U = np.zeros((len(conn), 3))
    U[:, 0] = h0
    return River(mesh=m, state=U, channel=channel, h_out=-slope * length - 1.5, q_peak=q_peak, t_ramp=t_ramp)


def cell_adjacency(m: Mesh):
    """Sparse symmetric cell adjacency through shared edges."""
    from scipy import sparse

    bnd = m.boundary_table()
    nb = bnd[0].T.astype(np.int64)                   # (nelt, 3), 1-based
    rows = np.repeat(np.arange(m.nelt), 3)
    cols = nb.reshape(-1) - 1
    keep = (nb.reshape(-1) > 0) & (cols != rows)
    A = sparse.csr_matrix((np.ones(keep.sum(), np.int8), (rows[keep], cols[keep])), shape=(m.nelt, m.nelt))
    return A


def rcm_order(m: Mesh) -> np.ndarray:
    """Reverse Cuthill-McKee cell order (new cell k = old cell perm[k])."""
    from scipy.sparse.csgraph import reverse_cuthill_mckee

    return np.asarray(reverse_cuthill_mckee(cell_adjacency(m), symmetric_mode=True), dtype=np.int64)


def random_order(m: Mesh, seed: int = 0) -> np.ndarray:
    return np.random.default_rng(seed).permutation(m.nelt)
