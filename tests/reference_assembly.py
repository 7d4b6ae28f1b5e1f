"""Sparse assembly of the reference's operators with scipy.sparse, following
src/Operators.jl literally (index loops, boundary branches, sparse products).
It is an independent restatement used only to pin the oracle's matrix-free
operators (not a test module)."""
import numpy as np
import scipy.sparse as sp


class Met:
    def __init__(self, o):
        for name in o.METRICS:
            setattr(self, name, o.metric(name))
        self.nx, self.ny = o.nx, o.ny


def _bcval(side):
    return float(getattr(side, "value", 0.0))


def _kind(side):
    return type(side).__name__


def scalar_operators(m, bc, px, py):
    """build_scalar_differential_operators, Operators.jl:189-356."""
    nx, ny = m.nx, m.ny
    N, Nfx, Nfy = nx * ny, (nx + 1) * ny, nx * (ny + 1)
    c = lambda i, j: i + j * nx
    fx = lambda i, j: i + j * (nx + 1)
    fy = lambda i, j: i + j * nx
    Gx, Gy = sp.lil_matrix((Nfx, N)), sp.lil_matrix((Nfy, N))
    gbx, gby = np.zeros(Nfx), np.zeros(Nfy)
    for j in range(ny):
        row = fx(0, j)
        if px:
            ih = 1.0 / m.dx_f[0]
            Gx[row, c(nx - 1, j)] += -ih
            Gx[row, c(0, j)] += ih
        elif _kind(bc.left) == "Dirichlet":
            d = m.dx_f[0] / 2.0
            Gx[row, c(0, j)] += 1.0 / d
            gbx[row] = -_bcval(bc.left) / d
        else:
            gbx[row] = -_bcval(bc.left)
        for i in range(1, nx):
            ih = 1.0 / m.dx_f[i]
            Gx[fx(i, j), c(i - 1, j)] += -ih
            Gx[fx(i, j), c(i, j)] += ih
        row = fx(nx, j)
        if px:
            ih = 1.0 / m.dx_f[-1]
            Gx[row, c(nx - 1, j)] += -ih
            Gx[row, c(0, j)] += ih
            gbx[row] = gbx[fx(0, j)]
        elif _kind(bc.right) == "Dirichlet":
            d = m.dx_f[-1] / 2.0
            Gx[row, c(nx - 1, j)] += -1.0 / d
            gbx[row] = _bcval(bc.right) / d
        else:
            gbx[row] = _bcval(bc.right)
    for i in range(nx):
        row = fy(i, 0)
        if py:
            ih = 1.0 / m.dy_f[0]
            Gy[row, c(i, ny - 1)] += -ih
            Gy[row, c(i, 0)] += ih
        elif _kind(bc.bottom) == "Dirichlet":
            d = m.dy_f[0] / 2.0
            Gy[row, c(i, 0)] += 1.0 / d
            gby[row] = -_bcval(bc.bottom) / d
        else:
            gby[row] = -_bcval(bc.bottom)
        for j in range(1, ny):
            ih = 1.0 / m.dy_f[j]
            Gy[fy(i, j), c(i, j - 1)] += -ih
            Gy[fy(i, j), c(i, j)] += ih
        row = fy(i, ny)
        if py:
            ih = 1.0 / m.dy_f[-1]
            Gy[row, c(i, ny - 1)] += -ih
            Gy[row, c(i, 0)] += ih
            gby[row] = gby[fy(i, 0)]
        elif _kind(bc.top) == "Dirichlet":
            d = m.dy_f[-1] / 2.0
            Gy[row, c(i, ny - 1)] += -1.0 / d
            gby[row] = _bcval(bc.top) / d
        else:
            gby[row] = _bcval(bc.top)
    Dx, Dy = sp.lil_matrix((N, Nfx)), sp.lil_matrix((N, Nfy))
    for j in range(ny):
        for i in range(nx):
            r = c(i, j)
            Dx[r, fx(i, j)] += -1.0 / m.dx_c[i]
            Dx[r, fx(i + 1, j)] += 1.0 / m.dx_c[i]
            Dy[r, fy(i, j)] += -1.0 / m.dy_c[j]
            Dy[r, fy(i, j + 1)] += 1.0 / m.dy_c[j]
    return Gx.tocsr(), Gy.tocsr(), Dx.tocsr(), Dy.tocsr(), gbx, gby


def pressure_laplacian(m, bc, px, py):
    """build_pressure_laplacian -> build_diffusion_operator with D = 1 (:361-390)."""
    Gx, Gy, Dx, Dy, gbx, gby = scalar_operators(m, bc, px, py)
    return (Dx @ Gx + Dy @ Gy).tocsr(), Dx @ gbx + Dy @ gby


def viscous_operator(m, mu, bcU, bcV, px, py):
    """build_velocity_gradient_operators (:848-924) + build_viscous_operator
    (:929-1083) with a constant viscosity field."""
    nx, ny = m.nx, m.ny
    N = nx * ny
    c = lambda i, j: i + j * nx
    ux, vy, uy, vx = (sp.lil_matrix((N, N)) for _ in range(4))
    ux_bc, vy_bc = np.zeros(N), np.zeros(N)
    for j in range(ny):
        for i in range(nx):
            r = c(i, j)
            ih = 1.0 / m.dx_c[i]
            ux[r, r] += -ih
            if i < nx - 1:
                ux[r, c(i + 1, j)] += ih
            elif px:
                ux[r, c(0, j)] += ih
            else:
                ux_bc[r] = _bcval(bcU.right) * ih
            ih = 1.0 / m.dy_c[j]
            vy[r, r] += -ih
            if j < ny - 1:
                vy[r, c(i, j + 1)] += ih
            elif py:
                vy[r, c(i, 0)] += ih
            else:
                vy_bc[r] = _bcval(bcV.top) * ih
            if j > 0:
                ih = 1.0 / m.dy_f[j]
                uy[r, c(i, j - 1)] += -ih
                uy[r, r] += ih
            elif py:
                ih = 1.0 / m.dy_f[0]
                uy[r, c(i, ny - 1)] += -ih
                uy[r, r] += ih
            if i > 0:
                ih = 1.0 / m.dx_f[i]
                vx[r, c(i - 1, j)] += -ih
                vx[r, r] += ih
            elif px:
                ih = 1.0 / m.dx_f[0]
                vx[r, c(nx - 1, j)] += -ih
                vx[r, r] += ih
    Z = sp.csr_matrix((N, N))
    E = sp.bmat([[ux, Z], [0.5 * uy, 0.5 * vx], [0.5 * uy, 0.5 * vx], [Z, vy]]).tocsr()
    E_bc = np.concatenate([ux_bc, np.zeros(N), np.zeros(N), vy_bc])
    w = np.full(4 * N, 2.0 * mu)
    dxx, dxy, dyx, dyy = (sp.lil_matrix((N, N)) for _ in range(4))
    for j in range(ny):
        for i in range(nx):
            r = c(i, j)
            ih = 1.0 / m.dx_f[i]
            dxx[r, r] += ih
            if i > 0:
                dxx[r, c(i - 1, j)] += -ih
            elif px:
                dxx[r, c(nx - 1, j)] += -ih
            ih = 1.0 / m.dy_c[j]
            dxy[r, r] += -ih
            if j < ny - 1:
                dxy[r, c(i, j + 1)] += ih
            elif py:
                dxy[r, c(i, 0)] += ih
            ih = 1.0 / m.dx_c[i]
            dyx[r, r] += -ih
            if i < nx - 1:
                dyx[r, c(i + 1, j)] += ih
            elif px:
                dyx[r, c(0, j)] += ih
            ih = 1.0 / m.dy_f[j]
            dyy[r, r] += ih
            if j > 0:
                dyy[r, c(i, j - 1)] += -ih
            elif py:
                dyy[r, c(i, ny - 1)] += -ih
    D = sp.bmat([[dxx, dxy, Z, Z], [Z, Z, dyx, dyy]]).tocsr()
    M = (D @ sp.diags(w) @ E).tolil()
    b = D @ (w * E_bc)
    if not py:
        db, dt = m.dy_f[0] / 2.0, m.dy_f[-1] / 2.0
        for i in range(nx):
            if _kind(bcU.bottom) == "Dirichlet":
                r = c(i, 0)
                wt = mu / (m.dy_c[0] * db)
                b[r] += wt * _bcval(bcU.bottom)
                M[r, r] -= wt
            if _kind(bcU.top) == "Dirichlet":
                r = c(i, ny - 1)
                wt = mu / (m.dy_c[ny - 1] * dt)
                b[r] += wt * _bcval(bcU.top)
                M[r, r] -= wt
    if not px:
        dl, dr = m.dx_f[0] / 2.0, m.dx_f[-1] / 2.0
        for j in range(ny):
            if _kind(bcV.left) == "Dirichlet":
                r = N + c(0, j)
                wt = mu / (m.dx_c[0] * dl)
                b[r] += wt * _bcval(bcV.left)
                M[r, r] -= wt
            if _kind(bcV.right) == "Dirichlet":
                r = N + c(nx - 1, j)
                wt = mu / (m.dx_c[nx - 1] * dr)
                b[r] += wt * _bcval(bcV.right)
                M[r, r] -= wt
    return M.tocsr(), b


def dirichlet_rows(nx, ny, bcU, bcV):
    """velocity_dirichlet_data, Operators.jl:1085-1110."""
    idx, val = [], []
    N = nx * ny
    if _kind(bcU.left) == "Dirichlet":
        for j in range(ny):
            idx.append(j * nx)
            val.append(_bcval(bcU.left))
    if _kind(bcV.bottom) == "Dirichlet":
        for i in range(nx):
            idx.append(N + i)
            val.append(_bcval(bcV.bottom))
    return idx, val
