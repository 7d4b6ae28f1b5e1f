"""Vectorised NumPy restatement of the reference's phase-field right-hand
sides, written from the Julia sources independently of the C oracle:

  upwind material flux        src/InterfaceTransport.jl:79-276
  CDI sharpening flux         src/InterfaceRegularization.jl:99-228
  ACDI (Jain)                 src/PhaseFieldMethods.jl:478-671
  UWCDI (Pirozzoli)           src/PhaseFieldMethods.jl:219-404, 675-867 (the
                              three normal reconstructions, the directional
                              right-hand sides and uwcdi_strang_step)
  DiffuseAreaCapillarity      src/Capillarity.jl:59-187
  face gradient / c2f / div   src/Operators.jl:110-159, 395-425 (gradient rows
                              from the sparse assembly in reference_assembly)

Arrays are indexed [i, j] with i along x: x faces (nx+1, ny), y faces
(nx, ny+1), cells (nx, ny).  Used by the oracle pins and the golden-fixture
script only (not a test module)."""
import numpy as np

from tests import reference_assembly as RA


class PhaseGeometry:
    """Metric arrays and boundary sides of one configuration."""

    def __init__(self, met, bc_phi, px, py):
        self.m, self.bc, self.px, self.py = met, bc_phi, px, py
        self.nx, self.ny = met.nx, met.ny
        self.ops = {}

    def gradient(self, f, bc=None):
        """scalar_gradient (Operators.jl:395-404): G f + g_bc on both face sets."""
        bc = self.bc if bc is None else bc
        key = id(bc)
        if key not in self.ops:
            self.ops[key] = RA.scalar_operators(self.m, bc, self.px, self.py)
        Gx, Gy, _, _, gbx, gby = self.ops[key]
        v = np.asarray(f, dtype=np.float64).ravel(order="F")
        gx = (Gx @ v + gbx).reshape((self.nx + 1, self.ny), order="F")
        gy = (Gy @ v + gby).reshape((self.nx, self.ny + 1), order="F")
        return gx, gy

    def c2f_x(self, f):
        """interpolate_center_to_x_faces (Operators.jl:110-134)."""
        m = self.m
        out = np.empty((self.nx + 1, self.ny))
        out[1:-1] = m.wL_c2fx[1:-1, None] * f[:-1] + m.wR_c2fx[1:-1, None] * f[1:]
        if self.px:
            out[0] = out[-1] = m.wL_c2fx[0] * f[-1] + m.wR_c2fx[0] * f[0]
        else:
            out[0], out[-1] = f[0], f[-1]
        return out

    def c2f_y(self, f):
        """interpolate_center_to_y_faces (Operators.jl:136-159)."""
        return self._swap().c2f_x(f.T).T

    def f2c(self, gx, gy):
        """Face-to-centre averages (PhaseFieldMethods.jl:219-232)."""
        m = self.m
        return (m.wL_f2cx[:, None] * gx[:-1] + m.wR_f2cx[:, None] * gx[1:],
                m.wB_f2cy[None, :] * gy[:, :-1] + m.wT_f2cy[None, :] * gy[:, 1:])

    def div(self, fx, fy):
        """divergence (Operators.jl:407-425)."""
        m = self.m
        return (fx[1:] - fx[:-1]) / m.dx_c[:, None] + (fy[:, 1:] - fy[:, :-1]) / m.dy_c[None, :]

    def _swap(self):
        """The same geometry with x and y exchanged (for the y-direction
        restatements written once along x)."""
        g = PhaseGeometry.__new__(PhaseGeometry)

        class M:
            pass
        s = M()
        m = self.m
        s.wL_c2fx, s.wR_c2fx = m.wB_c2fy, m.wT_c2fy
        g.m, g.px, g.nx, g.ny = s, self.py, self.ny, self.nx
        return g

    # boundary helpers ---------------------------------------------------
    def side_value(self, side, interior):
        """_phase_boundary_value (PhaseFieldMethods.jl:174-196):
        Dirichlet -> data, homogeneous Neumann -> the interior value."""
        kind = type(side).__name__
        if kind == "Dirichlet":
            return np.full_like(interior, float(side.value))
        if kind == "Neumann":
            assert abs(float(side.value)) <= 100 * np.finfo(float).eps
            return interior.copy()
        raise ValueError(kind)

    def zero_or_wrap(self, jx, jy):
        """_enforce_interface_flux_boundaries! (InterfaceRegularization.jl:150-170)."""
        jx, jy = jx.copy(), jy.copy()
        if self.px:
            jx[-1] = jx[0]
        else:
            jx[0] = jx[-1] = 0.0
        if self.py:
            jy[:, -1] = jy[:, 0]
        else:
            jy[:, 0] = jy[:, -1] = 0.0
        return jx, jy


def center_normals(g, f, bc, delta):
    """_center_normals (PhaseFieldMethods.jl:234-245)."""
    gx_f, gy_f = g.gradient(f, bc)
    gx, gy = g.f2c(gx_f, gy_f)
    mag = np.sqrt(gx ** 2 + gy ** 2 + delta ** 2)
    return gx / mag, gy / mag


# ---------------------------------------------------------------------------
# upwind material flux (InterfaceTransport.jl)

def _inflow_state(side, velocity, interior, inflow):
    """_x/_y_boundary_state (InterfaceTransport.jl:79-159)."""
    kind = type(side).__name__
    out = interior.copy()
    need = (velocity != 0.0) & inflow
    if np.any(need):
        if kind == "Dirichlet":
            out[need] = float(side.value)
        else:
            raise ValueError(f"{kind} phase boundary with advective inflow")
    return out


def upwind_fluxes(g, U, V, phi):
    """_x_advective_fluxes / _y_advective_fluxes (InterfaceTransport.jl:163-241)."""
    fx = np.empty((g.nx + 1, g.ny))
    fx[1:-1] = U[1:-1] * np.where(U[1:-1] >= 0.0, phi[:-1], phi[1:])
    if g.px:
        fx[0] = U[0] * np.where(U[0] >= 0.0, phi[-1], phi[0])
        fx[-1] = fx[0]
    else:
        b = _inflow_state(g.bc.left, U[0], phi[0], U[0] > 0.0)
        fx[0] = U[0] * np.where(U[0] >= 0.0, b, phi[0])
        b = _inflow_state(g.bc.right, U[-1], phi[-1], U[-1] < 0.0)
        fx[-1] = U[-1] * np.where(U[-1] >= 0.0, phi[-1], b)
    fy = np.empty((g.nx, g.ny + 1))
    fy[:, 1:-1] = V[:, 1:-1] * np.where(V[:, 1:-1] >= 0.0, phi[:, :-1], phi[:, 1:])
    if g.py:
        fy[:, 0] = V[:, 0] * np.where(V[:, 0] >= 0.0, phi[:, -1], phi[:, 0])
        fy[:, -1] = fy[:, 0]
    else:
        b = _inflow_state(g.bc.bottom, V[:, 0], phi[:, 0], V[:, 0] > 0.0)
        fy[:, 0] = V[:, 0] * np.where(V[:, 0] >= 0.0, b, phi[:, 0])
        b = _inflow_state(g.bc.top, V[:, -1], phi[:, -1], V[:, -1] < 0.0)
        fy[:, -1] = V[:, -1] * np.where(V[:, -1] >= 0.0, phi[:, -1], b)
    return fx, fy


def upwind_rhs(g, U, V, phi, S):
    """interface_explicit_rhs with zero interface flux (InterfaceTransport.jl:245-276)."""
    fx, fy = upwind_fluxes(g, U, V, phi)
    return -g.div(fx, fy) + phi * S


# ---------------------------------------------------------------------------
# CDI (InterfaceRegularization.jl)

def cdi_rhs(g, U, V, phi, S, inverse_peclet, epsilon, delta):
    """Upwind advection minus div of the conservative diffuse-interface flux
    j = Pe^-1 (phi(1-phi) n|_face - (epsilon/2) grad phi)."""
    gx_f, gy_f = g.gradient(phi)
    gx, gy = g.f2c(gx_f, gy_f)
    mag = np.sqrt(gx * gx + gy * gy + delta ** 2)
    amp = phi * (1.0 - phi)
    jx = inverse_peclet * (g.c2f_x(amp * gx / mag) - 0.5 * epsilon * gx_f)
    jy = inverse_peclet * (g.c2f_y(amp * gy / mag) - 0.5 * epsilon * gy_f)
    jx, jy = g.zero_or_wrap(jx, jy)
    return upwind_rhs(g, U, V, phi, S) - g.div(jx, jy)


# ---------------------------------------------------------------------------
# ACDI (PhaseFieldMethods.jl:478-671)

def _psi(p, epsilon, floor):
    p = np.clip(p, 0.0, 1.0)
    return epsilon * np.log((p + floor) / (1.0 - p + floor))


def _psi_bcs(bc, epsilon, floor):
    """_psi_boundary_setup: Dirichlet data mapped through psi, Neumann(0) kept."""
    import eulerianfsi_amd as fsi

    def side(s):
        kind = type(s).__name__
        if kind == "Dirichlet":
            return fsi.Dirichlet(float(_psi(np.float64(s.value), epsilon, floor)))
        if kind == "Neumann":
            return fsi.Neumann(0.0)
        return fsi.Periodic()
    return fsi.BoundaryConditionSetup(side(bc.left), side(bc.right), side(bc.bottom), side(bc.top))


def acdi_gamma(g, U, V, gamma, dynamic):
    """_acdi_effective_gamma: gamma * max |u| at cell centres when dynamic."""
    if not dynamic:
        return gamma
    uc, vc = g.f2c(U, V)
    return gamma * float(np.max(np.hypot(uc, vc)))


def acdi_rhs(g, U, V, phi, S, gamma, epsilon, floor=1e-100, delta=1e-14, dynamic=True):
    # convective flux with central face values, boundary data at walls
    fx = np.empty((g.nx + 1, g.ny))
    fy = np.empty((g.nx, g.ny + 1))
    m = g.m
    fx[1:-1] = U[1:-1] * (m.wL_c2fx[1:-1, None] * phi[:-1] + m.wR_c2fx[1:-1, None] * phi[1:])
    if g.px:
        fx[0] = U[0] * (m.wL_c2fx[0] * phi[-1] + m.wR_c2fx[0] * phi[0])
        fx[-1] = fx[0]
    else:
        fx[0] = U[0] * g.side_value(g.bc.left, phi[0])
        fx[-1] = U[-1] * g.side_value(g.bc.right, phi[-1])
    fy[:, 1:-1] = V[:, 1:-1] * (m.wB_c2fy[None, 1:-1] * phi[:, :-1] + m.wT_c2fy[None, 1:-1] * phi[:, 1:])
    if g.py:
        fy[:, 0] = V[:, 0] * (m.wB_c2fy[0] * phi[:, -1] + m.wT_c2fy[0] * phi[:, 0])
        fy[:, -1] = fy[:, 0]
    else:
        fy[:, 0] = V[:, 0] * g.side_value(g.bc.bottom, phi[:, 0])
        fy[:, -1] = V[:, -1] * g.side_value(g.bc.top, phi[:, -1])

    # regularisation: diffusion in phi, sharpening amplitude and normal in psi
    psi = _psi(phi, epsilon, floor)
    gpx, gpy = g.gradient(phi)
    nxc, nyc = center_normals(g, psi, _psi_bcs(g.bc, epsilon, floor), delta)
    ge = acdi_gamma(g, U, V, gamma, dynamic)
    ampx = 0.25 * (1.0 - np.tanh(g.c2f_x(psi) / (2.0 * epsilon)) ** 2)
    ampy = 0.25 * (1.0 - np.tanh(g.c2f_y(psi) / (2.0 * epsilon)) ** 2)
    Ax = ge * (epsilon * gpx - ampx * g.c2f_x(nxc))
    Ay = ge * (epsilon * gpy - ampy * g.c2f_y(nyc))
    Ax, Ay = g.zero_or_wrap(Ax, Ay)
    return -g.div(fx, fy) + g.div(Ax, Ay) + phi * S


# ---------------------------------------------------------------------------
# UWCDI (PhaseFieldMethods.jl:675-867)

def smoothed9(g, f):
    """The 9-point filter (PhaseFieldMethods.jl:364-383, Capillarity.jl:81-111):
    weights 1/4 (centre), 1/8 (edges), 1/16 (corners); an index past a wall
    is clamped, one across a periodic side wrapped."""
    nx, ny = f.shape
    ix = lambda off: (np.arange(nx) + off) % nx if g.px else np.clip(np.arange(nx) + off, 0, nx - 1)  # noqa: E731
    iy = lambda off: (np.arange(ny) + off) % ny if g.py else np.clip(np.arange(ny) + off, 0, ny - 1)  # noqa: E731
    F = lambda a, b: f[np.ix_(ix(a), iy(b))]  # noqa: E731
    return (0.25 * f + 0.125 * (F(1, 0) + F(-1, 0) + F(0, 1) + F(0, -1)) +
            0.0625 * (F(1, 1) + F(-1, 1) + F(1, -1) + F(-1, -1)))


def uwcdi_center_normals(g, phi, delta, method="standard"):
    """Centre normals of the UWCDI normal_reconstruction options
    (PhaseFieldMethods.jl:684-707): :standard, :youngs (:280-339),
    :smoothed9 (:341-404)."""
    if method == "smoothed9":
        for side in (g.bc.left, g.bc.right, g.bc.bottom, g.bc.top):
            kind = type(side).__name__
            assert kind == "Periodic" or (kind == "Neumann" and abs(float(side.value)) <= 100 * np.finfo(float).eps)
        return center_normals(g, smoothed9(g, phi), g.bc, delta)
    nxc, nyc = center_normals(g, phi, g.bc, delta)
    if method == "youngs":
        nx, ny = phi.shape
        hx, hy = g.m.dx_c[0], g.m.dy_c[0]
        im, ip = (np.arange(nx) - 1) % nx, (np.arange(nx) + 1) % nx
        jm, jp = (np.arange(ny) - 1) % ny, (np.arange(ny) + 1) % ny
        I, J = np.arange(nx), np.arange(ny)
        P = lambda a, b: phi[np.ix_(a, b)]  # noqa: E731
        gx = (P(ip, jm) + 2.0 * P(ip, J) + P(ip, jp) - P(im, jm) - 2.0 * P(im, J) - P(im, jp)) / (8.0 * hx)
        gy = (P(im, jp) + 2.0 * P(I, jp) + P(ip, jp) - P(im, jm) - 2.0 * P(I, jm) - P(ip, jm)) / (8.0 * hy)
        mag = np.sqrt(gx ** 2 + gy ** 2 + delta ** 2)
        keep = np.ones((nx, ny), dtype=bool)   # cells whose 3x3 stencil crosses a wall keep :standard
        if not g.px:
            keep[[0, -1], :] = False
        if not g.py:
            keep[:, [0, -1]] = False
        nxc = np.where(keep, gx / mag, nxc)
        nyc = np.where(keep, gy / mag, nyc)
    return nxc, nyc


def uwcdi_face_normals(g, phi, delta, method="standard"):
    """_normal_components_on_faces: centre normals interpolated to the faces
    and renormalised with the transverse component."""
    nxc, nyc = uwcdi_center_normals(g, phi, delta, method)
    nxf, nyx = g.c2f_x(nxc), g.c2f_x(nyc)
    nyf, nxy = g.c2f_y(nyc), g.c2f_y(nxc)
    nxf = nxf / np.sqrt(nxf ** 2 + nyx ** 2 + delta ** 2)
    nyf = nyf / np.sqrt(nxy ** 2 + nyf ** 2 + delta ** 2)
    if g.px:
        nxf[-1] = nxf[0]
    if g.py:
        nyf[:, -1] = nyf[:, 0]
    return nxf, nyf


def uwcdi_fluxes(g, U, V, phi, alpha=1.0, delta=1e-14, method="standard"):
    """_uwcdi_fluxes (PhaseFieldMethods.jl:675-766)."""
    m = g.m
    nxf, nyf = uwcdi_face_normals(g, phi, delta, method)
    # upwind face values
    up_x = np.empty((g.nx + 1, g.ny))
    up_x[1:-1] = np.where(U[1:-1] >= 0.0, phi[:-1], phi[1:])
    if g.px:
        up_x[0] = up_x[-1] = np.where(U[0] >= 0.0, phi[-1], phi[0])
    else:
        up_x[0] = np.where(U[0] >= 0.0, g.side_value(g.bc.left, phi[0]), phi[0])
        up_x[-1] = np.where(U[-1] >= 0.0, phi[-1], g.side_value(g.bc.right, phi[-1]))
    up_y = np.empty((g.nx, g.ny + 1))
    up_y[:, 1:-1] = np.where(V[:, 1:-1] >= 0.0, phi[:, :-1], phi[:, 1:])
    if g.py:
        up_y[:, 0] = up_y[:, -1] = np.where(V[:, 0] >= 0.0, phi[:, -1], phi[:, 0])
    else:
        up_y[:, 0] = np.where(V[:, 0] >= 0.0, g.side_value(g.bc.bottom, phi[:, 0]), phi[:, 0])
        up_y[:, -1] = np.where(V[:, -1] >= 0.0, phi[:, -1], g.side_value(g.bc.top, phi[:, -1]))
    # multidimensional correction of alpha (section 2.3 of Pirozzoli's paper)
    ax = alpha * np.maximum(1.0, m.dy_c[None, :] / m.dx_f[:, None])
    ay = alpha * np.maximum(1.0, m.dx_c[:, None] / m.dy_f[None, :])
    fx = up_x * (U + np.abs(U) / ax * nxf * (1.0 - up_x))
    fy = up_y * (V + np.abs(V) / ay * nyf * (1.0 - up_y))
    if g.px:
        fx[-1] = fx[0]
    if g.py:
        fy[:, -1] = fy[:, 0]
    return fx, fy


def uwcdi_rhs(g, U, V, phi, S, alpha=1.0, delta=1e-14, method="standard"):
    fx, fy = uwcdi_fluxes(g, U, V, phi, alpha, delta, method)
    return -g.div(fx, fy) + phi * S


def uwcdi_directional_rhs(g, model, direction, U, V, phi):
    """uwcdi_directional_rhs (PhaseFieldMethods.jl:768-791): minus the
    divergence of one flux component, the other set to zero."""
    fx, fy = uwcdi_fluxes(g, U, V, phi, model.alpha, model.delta_normal, model.normal_reconstruction)
    if direction == "x":
        return -g.div(fx, np.zeros((g.nx, g.ny + 1)))
    return -g.div(np.zeros((g.nx + 1, g.ny)), fy)


def uwcdi_strang_step(g, model, U, V, phi, dt, first_direction="x"):
    """uwcdi_strang_step (PhaseFieldMethods.jl:793-850): Euler sweeps a (dt/2),
    b (dt), a (dt/2) with (a, b) = (x, y) or (y, x)."""
    a, b = ("x", "y") if first_direction == "x" else ("y", "x")
    q = np.array(phi, dtype=np.float64)
    for d, h in ((a, 0.5 * dt), (b, dt), (a, 0.5 * dt)):
        if h != 0.0:
            q = q + h * uwcdi_directional_rhs(g, model, d, U, V, q)
    return q


def capillary_force(g, phi, coefficient, delta_normal=1e-6, smoothing_passes=0):
    """capillary_force(::DiffuseAreaCapillarity) (Capillarity.jl:134-187):
    the normal of the (optionally filtered) field, kappa = -div(n), and
    coefficient kappa grad(phi) on both face sets."""
    geo = np.array(phi, dtype=np.float64)
    for _ in range(smoothing_passes):
        geo = smoothed9(g, geo)
    ggx, ggy = g.gradient(geo)
    gx, gy = g.gradient(phi)
    gxc, gyc = g.f2c(ggx, ggy)
    den = np.sqrt(gxc * gxc + gyc * gyc + delta_normal ** 2)
    kappa = -g.div(g.c2f_x(gxc / den), g.c2f_y(gyc / den))
    return coefficient * g.c2f_x(kappa) * gx, coefficient * g.c2f_y(kappa) * gy


def phase_rhs(g, model, U, V, phi, S):
    """Dispatch on the closure types of eulerianfsi_amd.closures."""
    name = type(model).__name__
    if name == "LegacySplitPhaseFieldMethod":
        flux = model.interface_flux
        if type(flux).__name__ == "NoInterfaceFlux":
            return upwind_rhs(g, U, V, phi, S)
        return cdi_rhs(g, U, V, phi, S, flux.inverse_peclet, flux.epsilon, flux.delta_normal)
    if name == "NoPhaseFieldRegularization":
        return upwind_rhs(g, U, V, phi, S)
    if name == "CurrentCDIPhaseField":
        f = model.flux
        return cdi_rhs(g, U, V, phi, S, f.inverse_peclet, f.epsilon, f.delta_normal)
    if name == "JainACDIPhaseField":
        return acdi_rhs(g, U, V, phi, S, model.gamma, model.epsilon, model.psi_floor, model.delta_normal,
                        model.dynamic_gamma)
    if name == "PirozzoliUWCDIPhaseField":
        return uwcdi_rhs(g, U, V, phi, S, model.alpha, model.delta_normal, model.normal_reconstruction)
    raise ValueError(name)
