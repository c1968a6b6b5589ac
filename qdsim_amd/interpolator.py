"""FE field evaluation at arbitrary points on the GPU.

Drop-in for the reference's FEInterpolator binding (backend/src/bindings.cpp:779-808,
implementation backend/src/fe_interpolator.cpp:54-613): ``interpolate(x, y, field)``,
``interpolate_with_gradient(x, y, field) -> (value, grad_x, grad_y)``,
``find_element(x, y)``, ``compute_barycentric_coordinates(x, y, vertices) ->
(inside, lambda)``, ``evaluate_shape_functions(lambda)`` and
``evaluate_shape_function_gradients(lambda, vertices)``.  Coordinates are in the
units of the mesh (``SimpleMesh.get_nodes``), gradients in field units per
coordinate unit.

The point queries run in the qds_interpolate_points kernel; ``interpolate_points``
is the batched form (device or host arrays of points) that the per-point methods
call with one point.  Points outside the mesh follow the reference: the value of
the nearest node, and for the gradient the inverse-distance estimate over the
nodes sharing an element with it.  The barycentric and shape-function helpers are
the closed forms of fe_interpolator.cpp:433-455, :474-518 and :542-613 evaluated
on the host for a single point, as the binding exposes them.
"""
import numpy as np
import torch

from . import _lib
from .device import device, ptr, stream_ptr


class FEInterpolator:
    def __init__(self, mesh):
        self.mesh = mesh
        self.order = int(getattr(mesh, "element_order", 1))
        dev = device()
        x, y = mesh.get_nodes()
        conn = mesh.get_elements()
        self.cx, self.cy = mesh.nx - 1, mesh.ny - 1
        self.dx = mesh.Lx / self.cx
        self.dy = mesh.Ly / self.cy
        # lowest node id at every refined-lattice position (unused orphan DOFs skipped)
        used = np.zeros(mesh.num_nodes, dtype=bool)
        used[conn.ravel()] = True
        lidx = np.asarray(mesh.lattice_index())
        NY, NX = mesh.lattice_shape
        lnode = np.full(NX * NY, np.iinfo(np.int32).max, dtype=np.int64)
        ids = np.nonzero(used)[0]
        np.minimum.at(lnode, lidx[ids], ids)
        if np.any(lnode == np.iinfo(np.int32).max):
            raise ValueError("mesh DOFs do not cover the refined lattice")
        self._x = torch.from_numpy(np.ascontiguousarray(x, dtype=np.float64)).to(dev)
        self._y = torch.from_numpy(np.ascontiguousarray(y, dtype=np.float64)).to(dev)
        self._conn = torch.from_numpy(np.ascontiguousarray(conn, dtype=np.int32)).to(dev)
        self._lnode = torch.from_numpy(lnode.astype(np.int32)).to(dev)
        self._hx, self._hy = x, y

    # -- batched device evaluation -------------------------------------------------
    def interpolate_points(self, px, py, field, gradient=False, return_element=False):
        """Values (and gradients) of the nodal field(s) at the points (px, py).

        field: (num_nodes,) or (num_nodes, nf).  Inputs may be numpy arrays or device
        tensors; outputs are FP64 device tensors of shape (n,) or (n, nf)."""
        dev = self._x.device
        px = torch.as_tensor(px, dtype=torch.float64).to(dev).reshape(-1).contiguous()
        py = torch.as_tensor(py, dtype=torch.float64).to(dev).reshape(-1).contiguous()
        f = torch.as_tensor(field, dtype=torch.float64).to(dev)
        multi = f.dim() == 2
        f = f.reshape(f.shape[0], -1).contiguous()
        nf = f.shape[1]
        if f.shape[0] < self.mesh.num_nodes:
            raise ValueError("field must have one value per mesh node")
        n = px.numel()
        shape = (n, nf) if multi else (n,)
        val = torch.empty(shape, dtype=torch.float64, device=dev)
        gx = torch.empty(shape, dtype=torch.float64, device=dev) if gradient else None
        gy = torch.empty(shape, dtype=torch.float64, device=dev) if gradient else None
        el = torch.empty(n, dtype=torch.int32, device=dev) if return_element else None
        _lib.check(_lib.lib.qds_interpolate_points(self.order, self.cx, self.cy, self.dx, self.dy, ptr(self._conn),
                                                   ptr(self._x), ptr(self._y), ptr(self._lnode), ptr(f), nf, n,
                                                   ptr(px), ptr(py), ptr(val), ptr(gx), ptr(gy), ptr(el),
                                                   stream_ptr()), "qds_interpolate_points")
        out = (val, gx, gy) if gradient else (val,)
        if return_element:
            out = out + (el,)
        return out if len(out) > 1 else out[0]

    # -- reference per-point API ---------------------------------------------------
    def interpolate(self, x, y, field):
        return float(self.interpolate_points([x], [y], field)[0].item())

    def interpolate_with_gradient(self, x, y, field):
        v, gx, gy = self.interpolate_points([x], [y], field, gradient=True)
        return float(v.item()), float(gx.item()), float(gy.item())

    def find_element(self, x, y):
        f = torch.zeros(self.mesh.num_nodes, dtype=torch.float64, device=self._x.device)
        _, el = self.interpolate_points([x], [y], f, return_element=True)
        return int(el.item())

    @staticmethod
    def compute_barycentric_coordinates(x, y, vertices):
        (x0, y0), (x1, y1), (x2, y2) = [tuple(map(float, v)) for v in vertices]
        area = 0.5 * ((x1 - x0) * (y2 - y0) - (x2 - x0) * (y1 - y0))
        lam = [0.5 * ((x1 * y2 - x2 * y1) + (y1 - y2) * x + (x2 - x1) * y) / area,
               0.5 * ((x2 * y0 - x0 * y2) + (y2 - y0) * x + (x0 - x2) * y) / area,
               0.5 * ((x0 * y1 - x1 * y0) + (y0 - y1) * x + (x1 - x0) * y) / area]
        return all(v >= -1e-10 for v in lam), lam

    def evaluate_shape_functions(self, lam):
        l0, l1, l2 = (float(v) for v in lam)
        if self.order == 1:
            return [l0, l1, l2]
        if self.order == 2:
            return [l0 * (2 * l0 - 1), l1 * (2 * l1 - 1), l2 * (2 * l2 - 1), 4 * l0 * l1, 4 * l1 * l2, 4 * l2 * l0]
        c = lambda a: 0.5 * a * (3 * a - 1) * (3 * a - 2)
        e = lambda a, b: 4.5 * a * b * (3 * a - 1)
        return [c(l0), c(l1), c(l2), e(l0, l1), e(l1, l0), e(l1, l2), e(l2, l1), e(l2, l0), e(l0, l2),
                27 * l0 * l1 * l2]

    def evaluate_shape_function_gradients(self, lam, vertices):
        """[(dN/dx, dN/dy)] per node: d/d lambda of the shape functions times grad lambda."""
        (x0, y0), (x1, y1), (x2, y2) = [tuple(map(float, v)) for v in vertices[:3]]
        area = 0.5 * ((x1 - x0) * (y2 - y0) - (x2 - x0) * (y1 - y0))
        g = [np.array([0.5 * (y1 - y2) / area, 0.5 * (x2 - x1) / area]),
             np.array([0.5 * (y2 - y0) / area, 0.5 * (x0 - x2) / area]),
             np.array([0.5 * (y0 - y1) / area, 0.5 * (x1 - x0) / area])]
        l = [float(v) for v in lam]
        if self.order == 1:
            return [tuple(v) for v in g]
        if self.order == 2:
            out = [g[a] * (4 * l[a] - 1) for a in range(3)]
            out += [4 * (g[i] * l[j] + g[j] * l[i]) for i, j in ((0, 1), (1, 2), (2, 0))]
            return [tuple(v) for v in out]
        out = [g[a] * (0.5 * (27 * l[a] ** 2 - 18 * l[a] + 2)) for a in range(3)]
        for i, j in ((0, 1), (1, 2), (2, 0)):
            # node nearer i: 4.5 li lj (3 li - 1); node nearer j: 4.5 li lj (3 lj - 1)
            out.append(4.5 * (g[i] * l[j] * (6 * l[i] - 1) + g[j] * l[i] * (3 * l[i] - 1)))
            out.append(4.5 * (g[j] * l[i] * (6 * l[j] - 1) + g[i] * l[j] * (3 * l[j] - 1)))
        out.append(27 * (g[0] * l[1] * l[2] + g[1] * l[0] * l[2] + g[2] * l[0] * l[1]))
        return [tuple(v) for v in out]


def create_interpolator(mesh):
    return FEInterpolator(mesh)
