"""RectilinearGrid (src/Grids/rectilinear_grid.jl:3-65, src/Grids/grid_generation.jl) — host description.

Uniform x and y; z is uniform (a tuple) or stretched (a vector of Nz+1 faces or a function of
the 1-based face index).  Topologies: Periodic or Flat in x and y, Bounded / Periodic / Flat in z.
"""
import numpy as np


class Periodic:
    code = 0


class Bounded:
    code = 1


class Flat:
    code = 2


class MI355X:
    """The architecture: one MI355X (gfx950) GPU driven through the C ABI."""

    def __repr__(self):
        return "MI355X()"


def _topo(t):
    if isinstance(t, type) and t in (Periodic, Bounded, Flat):
        return t
    if isinstance(t, str):
        return {"periodic": Periodic, "bounded": Bounded, "flat": Flat}[t.lower()]
    raise ValueError(f"unknown topology {t!r}")


class RectilinearGrid:
    def __init__(self, architecture=None, float_type=np.float64, *, size, x=None, y=None, z=None,
                 extent=None, topology=(Periodic, Periodic, Bounded), halo=None):
        self.architecture = architecture if architecture is not None else MI355X()
        self.float_type = np.dtype(float_type).type
        if self.float_type not in (np.float32, np.float64):
            raise ValueError("float_type must be float32 or float64")
        self.topology = tuple(_topo(t) for t in topology)
        size = tuple(size) if np.iterable(size) else (size,)
        nonflat = [d for d in range(3) if self.topology[d] is not Flat]
        if len(size) != len(nonflat):
            raise ValueError("size must list one entry per non-Flat dimension")
        N = [1, 1, 1]
        for d, n in zip(nonflat, size):
            N[d] = int(n)
        self.N = tuple(N)
        if halo is None:
            halo = tuple(3 for _ in nonflat)
        halo = tuple(halo) if np.iterable(halo) else (halo,)
        H = [0, 0, 0]
        for d, h in zip(nonflat, halo):
            H[d] = int(h)
        self.H = tuple(H)
        coords = [x, y, z]
        if extent is not None:
            extent = tuple(extent)
            e = list(extent)
            for q, d in enumerate(nonflat):
                L = e[q]
                coords[d] = (0.0, L) if d < 2 else (-L, 0.0)
        self.origin = [0.0, 0.0, 0.0]
        self.extent = [1.0, 1.0, 1.0]
        self.z_faces = None
        for d in range(3):
            if self.topology[d] is Flat:
                continue
            c = coords[d]
            if c is None:
                raise ValueError(f"coordinate {'xyz'[d]} is required")
            if d == 2 and (callable(c) or (np.iterable(c) and len(c) == N[2] + 1 and N[2] + 1 != 2)):
                faces = np.array([c(k) for k in range(1, N[2] + 2)] if callable(c) else c, dtype=np.float64)
                if np.any(np.diff(faces) <= 0):
                    raise ValueError("the elements of z must be increasing")
                self.z_faces = faces
                self.origin[2] = float(faces[0])
                self.extent[2] = float(faces[-1] - faces[0])
                continue
            if callable(c) or (np.iterable(c) and len(c) != 2):
                raise ValueError("only z may be stretched on the MI355X path")
            lo, hi = float(c[0]), float(c[1])
            if hi <= lo:
                raise ValueError(f"{'xyz'[d]} must be an increasing interval")
            self.origin[d] = lo
            self.extent[d] = hi - lo

    # sizes
    Nx = property(lambda s: s.N[0])
    Ny = property(lambda s: s.N[1])
    Nz = property(lambda s: s.N[2])
    Lx = property(lambda s: s.extent[0])
    Ly = property(lambda s: s.extent[1])
    Lz = property(lambda s: s.extent[2])

    @property
    def stretched(self):
        return self.z_faces is not None

    # node positions (float64), 1-based interior
    def xnodes(self, loc="c"):
        if self.topology[0] is Flat:
            return np.zeros(1)
        dx = self.extent[0] / self.N[0]
        i = np.arange(self.N[0])
        return self.origin[0] + (i + (0.5 if loc == "c" else 0.0)) * dx

    def ynodes(self, loc="c"):
        if self.topology[1] is Flat:
            return np.zeros(1)
        dy = self.extent[1] / self.N[1]
        j = np.arange(self.N[1])
        return self.origin[1] + (j + (0.5 if loc == "c" else 0.0)) * dy

    def znodes(self, loc="c"):
        if self.topology[2] is Flat:
            return np.zeros(1)
        if self.z_faces is not None:
            F = self.z_faces
        else:
            F = self.origin[2] + np.arange(self.N[2] + 1) * (self.extent[2] / self.N[2])
        if loc == "c":
            return 0.5 * (F[1:] + F[:-1])
        return F if self.topology[2] is Bounded else F[:-1]

    def __repr__(self):
        t = ", ".join(t.__name__ for t in self.topology)
        return (f"{self.N[0]}×{self.N[1]}×{self.N[2]} RectilinearGrid{{{np.dtype(self.float_type).name}, {t}}} "
                f"on {self.architecture!r} with {self.H[0]}×{self.H[1]}×{self.H[2]} halo")


class GridFittedBottom:
    """GridFittedBottom(bottom_height) (src/ImmersedBoundaries/grid_fitted_bottom.jl:17-20): cells
    whose centre lies at or below the bottom height are immersed (CenterImmersedCondition).
    bottom_height: a function (x, y) → z_b or an (Nx, Ny) array at the cell centres."""

    def __init__(self, bottom_height):
        self.bottom_height = bottom_height

    def heights(self, grid):
        h = self.bottom_height
        if callable(h):
            X, Y = np.meshgrid(grid.xnodes("c"), grid.ynodes("c"), indexing="ij")
            return np.asarray(h(X, Y), dtype=np.float64) * np.ones((grid.Nx, grid.Ny))
        h = np.asarray(h, dtype=np.float64)
        if h.shape != (grid.Nx, grid.Ny):
            raise ValueError(f"bottom height must have shape {(grid.Nx, grid.Ny)}")
        return h

    def __repr__(self):
        return f"GridFittedBottom({self.bottom_height!r})"


class ImmersedBoundaryGrid:
    """ImmersedBoundaryGrid(grid, immersed_boundary) (src/ImmersedBoundaries/immersed_boundary_grid.jl):
    the underlying RectilinearGrid with a GridFittedBottom.  Every grid attribute is the underlying
    grid's."""

    def __init__(self, grid, immersed_boundary):
        if not isinstance(grid, RectilinearGrid):
            raise TypeError("the underlying grid must be a RectilinearGrid")
        if not isinstance(immersed_boundary, GridFittedBottom):
            raise TypeError("only GridFittedBottom immersed boundaries are supported")
        if grid.topology[2] is not Bounded or any(t is Flat for t in grid.topology):
            raise ValueError("a GridFittedBottom needs a three-dimensional grid with a Bounded z")
        self.underlying_grid = grid
        self.immersed_boundary = immersed_boundary

    def __getattr__(self, name):
        return getattr(self.__dict__["underlying_grid"], name)

    def bottom_heights(self):
        return self.immersed_boundary.heights(self.underlying_grid)

    def __repr__(self):
        return f"ImmersedBoundaryGrid({self.underlying_grid!r}, {self.immersed_boundary!r})"
