"""Host-side grid of setup.F90:92-247 (uniform, or with the tanh stretch of
setup.F90:256-285 on any axis).

The grid stays on the host as in the reference; it produces the 1-D metric
arrays (xb, xc, dxb, dxc per axis) the device library receives at init.
"""
import numpy as np

from .abi import BC_PERIODIC


def tanh_stretch(xb, n, length, new_length):
    """Stretch node coordinates xb(-2:n+2) (indexed [i+2]) of a uniform axis
    in place, as stretch_x does (setup.F90:256-285): the cell widths grow by
    f (1 + tanh((|x| - L) / width)) dx with L = length / 1.5, width =
    length / 10 and f chosen so that the axis spans new_length."""
    L = length / 1.5
    width = length / 10.0
    f = (new_length - length) / (length - L) / 2.0
    dx = length / float(n)
    dxnew = dx + f * (1.0 + np.tanh((np.abs(xb) - L) / width)) * dx
    for i in range(1, n + 3):
        xb[i + 2] = xb[i + 1] + dxnew[i + 2]


def axis_metrics(n, lo, hi, periodic, stretch=None, widths=None):
    """Return (xb(-2:n+2), xc(-1:n+2), dxb(-1:n+2), dxc(-1:n+1)) for one axis,
    following setup.F90:106-149 (one MPI rank, so local == global).
    stretch = the stretched length of the axis, or None for a uniform one.
    widths = the n cell widths of cells 1..n (nodes from lo), or None; the
    C-ABI takes any 1-D metrics, so a grid need not come from the
    reference's uniform or tanh-stretched forms."""
    d = (hi - lo) / float(n)
    xb = np.array([lo + float(i) * d for i in range(-2, n + 3)])  # index i+2
    if stretch is not None:
        tanh_stretch(xb, n, hi - lo, stretch)
    if widths is not None:
        w = np.asarray(widths, dtype=float)
        assert w.shape == (n,)
        xb[2] = lo
        for i in range(1, n + 1):
            xb[i + 2] = xb[i + 1] + w[i - 1]
    B = lambda i: xb[i + 2]
    if periodic:
        xb[n + 1 + 2] = B(n) + (B(1) - B(0))
        xb[n + 2 + 2] = B(n) + (B(2) - B(0))
        xb[-1 + 2] = B(0) - (B(n) - B(n - 1))
        xb[-2 + 2] = B(0) - (B(n) - B(n - 2))
    else:
        xb[n + 1 + 2] = 2.0 * B(n) - B(n - 1)
        xb[n + 2 + 2] = 2.0 * B(n) - B(n - 2)
        xb[-1 + 2] = 2.0 * B(0) - B(1)
        xb[-2 + 2] = 2.0 * B(0) - B(2)
    xc = np.empty(n + 4)   # -1..n+2
    dxb = np.empty(n + 4)  # -1..n+2
    for i in range(-1, n + 3):
        xc[i + 1] = 0.5 * (B(i - 1) + B(i))
        dxb[i + 1] = B(i) - B(i - 1)
    dxc = np.empty(n + 3)  # -1..n+1
    for i in range(-1, n + 2):
        dxc[i + 1] = xc[i + 2] - xc[i + 1]
    return xb, xc, dxb, dxc


class Grid:
    """Sizes, boundary codes, extents and metrics of one run."""

    def __init__(self, nx, ny, nz, extent, bcs, stretch=(None, None, None), widths=(None, None, None)):
        self.nx, self.ny, self.nz = nx, ny, nz
        self.n = (nx, ny, nz)
        self.extent = extent            # ((x_min,x_max),(y_min,y_max),(z_min,z_max))
        self.bcs = tuple(bcs)           # (xbc_min, xbc_max, ybc_min, ybc_max, zbc_min, zbc_max)
        self.stretch = tuple(stretch)   # stretched length per axis, or None
        self.widths = tuple(widths)     # explicit cell widths per axis, or None
        self.xb, self.xc, self.dxb, self.dxc = [], [], [], []
        for a in range(3):
            per = self.bcs[2 * a] == BC_PERIODIC and self.bcs[2 * a + 1] == BC_PERIODIC
            xb, xc, db, dc = axis_metrics(self.n[a], extent[a][0], extent[a][1], per, self.stretch[a],
                                          self.widths[a])
            self.xb.append(xb); self.xc.append(xc); self.dxb.append(db); self.dxc.append(dc)
        self.grav = np.zeros(nz + 4)

    # coordinates with their Fortran index offsets
    def cb(self, a, i):
        return self.xb[a][i + 2]

    def cc(self, a, i):
        return self.xc[a][i + 1]

    def node_coords(self, a):
        """xb(-2:n+2) as an array indexed [i+2]."""
        return self.xb[a]

    def cell_coords(self, a):
        """xc(-1:n+2) as an array indexed [i+1]."""
        return self.xc[a]


def grid_struct(grid):
    """LareGrid for the C-ABI.  Keeps the metric arrays alive on the struct."""
    from .abi import LareGrid, dptr
    g = LareGrid()
    g.nx, g.ny, g.nz = grid.n
    (g.xbc_min, g.xbc_max, g.ybc_min, g.ybc_max, g.zbc_min, g.zbc_max) = grid.bcs
    keep = [np.ascontiguousarray(a) for a in (grid.dxb[0], grid.dxc[0], grid.dxb[1], grid.dxc[1],
                                              grid.dxb[2], grid.dxc[2], grid.grav)]
    g.dxb, g.dxc, g.dyb, g.dyc, g.dzb, g.dzc, g.grav = [dptr(a) for a in keep]
    g._keep = keep
    return g
