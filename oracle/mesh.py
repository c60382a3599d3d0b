"""Spectral spherical mesh: tables, transforms and operators (oracle).

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Restates ``src/Atmos_Spectral/Spectral_Spherical_Mesh.jl`` (abbreviated SSM):

==============================  ==========================================
oracle                          reference
==============================  ==========================================
``SpectralSphericalMesh``       struct + constructor SSM:6-139
``trans_spherical_to_grid``     ``Trans_Spherical_To_Grid!`` SSM:143-217
``trans_grid_to_spherical``     ``Trans_Grid_To_Spherical!`` SSM:221-293
``trans_grid_to_fourier``       ``Trans_Grid_To_Fourier!`` SSM:295-330
``divide_by_cos`` / ``multiply_by_cos``   SSM:333-346
``vor_div_from_grid_uv``        ``Vor_Div_From_Grid_UV!`` SSM:349-393
``compute_alpha_operator``      ``Compute_Alpha_Operator!`` SSM:396-459
``uv_grid_from_vor_div``        ``UV_Grid_From_Vor_Div!`` SSM:462-482
``compute_ucos_vcos_from_vor_div``  SSM:485-580
``apply_laplacian``             ``Apply_Laplacian!`` SSM:583-621
``compute_gradient_cos``        ``Compute_Gradient_Cos!`` SSM:625-692
``add_horizontal_advection``    ``Add_Horizontal_Advection!`` SSM:694-718
``compute_gradients``           ``Compute_Gradients!`` SSM:721-741
``area_weighted_global_mean``   ``Area_Weighted_Global_Mean`` SSM:746-760
==============================  ==========================================

The FFT stage uses ``numpy.fft`` (pocketfft) where the reference uses FFTW;
both compute the DFT definition, so only rounding differs.  Pins:
``tests/test_oracle_mesh.py`` (single-harmonic coefficients, round trips,
the analytic velocity/vorticity fields of SSM:804-1008, Laplacian closed form).
"""
import numpy as np

from .gauss_legendre import compute_gaussian, compute_legendre


class SpectralSphericalMesh:
    """Mesh tables (SSM:69-139).  Argument order follows the reference:
    ``(num_fourier, num_spherical, nθ, nλ, nd, radius)`` (nθ before nλ)."""

    def __init__(self, num_fourier, num_spherical, nlat, nlon, nd, radius):
        N, ns = num_fourier, num_spherical
        self.num_fourier, self.num_spherical = N, ns
        self.nlat, self.nlon, self.nd = nlat, nlon, nd
        self.radius = float(radius)
        sin_theta, wts = compute_gaussian(nlat)
        qnm, dqnm = compute_legendre(N, ns, sin_theta, nlat)
        self.sin_theta, self.wts = sin_theta, wts
        self.cos_theta = np.sqrt(1.0 - sin_theta ** 2)
        self.qnm, self.dqnm = qnm, dqnm
        # SSM:78-92 grid coordinates and cell edges
        self.lam_c = np.linspace(0.0, 2 * np.pi, nlon + 1)[:nlon]
        self.theta_c = np.arcsin(sin_theta)
        dlam = 2 * np.pi / nlon
        self.lam_e = np.linspace(-0.5, nlon - 0.5, nlon + 1) * dlam
        theta_e = np.zeros(nlat + 1)
        theta_e[0] = -0.5 * np.pi
        acc = 0.0
        for j in range(nlat):
            acc += wts[j]
            theta_e[j + 1] = np.arcsin(acc - 1.0)
        theta_e[nlat] = 0.5 * np.pi
        self.theta_e = theta_e
        # SSM:96-99
        self.qwg = qnm * wts[None, None, :]
        # SSM:101-106
        eps = np.zeros((N + 1, ns + 1))
        for m in range(N + 1):
            for n in range(m, ns + 1):
                eps[m, n] = np.sqrt((n * n - m * m) / (4.0 * n * n - 1.0))
        self.epsilon = eps
        c2 = self.cos_theta ** 2
        # SSM:396-422 (degrees m..num_spherical)
        a = np.zeros((N + 1, ns + 1, nlat))
        b = np.zeros((N + 1, ns + 1, nlat))
        for m in range(N + 1):
            for n in range(m, ns + 1):
                a[m, n, :] = wts * qnm[m, n, :] / c2 * m
                b[m, n, :] = -wts * dqnm[m, n, :]
        self.coef_alp_a, self.coef_alp_b = a, b
        # SSM:485-515
        uvm = np.zeros((N + 1, ns + 1))
        uvc = np.zeros((N + 1, ns + 1))
        uvp = np.zeros((N + 1, ns + 1))
        r = self.radius
        for m in range(N + 1):
            for n in range(m, ns + 1):
                if n > 0:
                    uvc[m, n] = -r * m / (n * (n + 1))
                    uvm[m, n] = r * eps[m, n] / n
            for n in range(m, ns):
                uvp[m, n] = -r * eps[m, n + 1] / (n + 1)
        self.coef_uvm, self.coef_uvc, self.coef_uvp = uvm, uvc, uvp
        # SSM:625-652
        dlam_c = np.zeros((N + 1, ns + 1))
        dthm = np.zeros((N + 1, ns + 1))
        dthp = np.zeros((N + 1, ns + 1))
        for m in range(N + 1):
            for n in range(m, ns + 1):
                dlam_c[m, n] = m / r
                dthm[m, n] = -(n - 1) * eps[m, n] / r
            for n in range(m, ns):
                dthp[m, n] = (n + 2) * eps[m, n + 1] / r
        self.coef_dlam, self.coef_dthm, self.coef_dthp = dlam_c, dthm, dthp
        # SSM:583-613
        eig = np.zeros((N + 1, ns + 1))
        wn = np.zeros((N + 1, ns + 1), dtype=np.int64)
        for m in range(N + 1):
            for n in range(m, ns + 1):
                eig[m, n] = -n * (n + 1) / r ** 2
                wn[m, n] = n
        self.laplacian_eig, self.wave_numbers = eig, wn

    # ------------------------------------------------------------------ shapes
    def spec_zeros(self, nlev=None):
        nlev = self.nd if nlev is None else nlev
        return np.zeros((nlev, self.num_spherical + 1, self.num_fourier + 1), np.complex128)

    def grid_zeros(self, nlev=None):
        nlev = self.nd if nlev is None else nlev
        return np.zeros((nlev, self.nlat, self.nlon), np.float64)

    def _check_levels(self, nlev):
        # SSM:175,255,370,466,707,724
        assert nlev == self.nd or nlev == 1

    # ----------------------------------------------------------- transforms
    def trans_spherical_to_grid(self, snm, pfield):
        """SSM:143-217.  Sums degrees m..num_spherical (the N+1 row too,
        SSM:196), unfolds even/odd halves, halves m=0, 2nλ Re(ifft)."""
        self._check_levels(snm.shape[0])
        N, ns = self.num_fourier, self.num_spherical
        nlat, nlon = self.nlat, self.nlon
        assert nlat % 2 == 0
        h = nlat // 2
        nlev = snm.shape[0]
        g = np.zeros((nlev, nlat, nlon), np.complex128)
        for m in range(N + 1):
            n = np.arange(m, ns + 1)
            ev = n[(n - m) % 2 == 0]
            od = n[(n - m) % 2 == 1]
            q = self.qnm[m, :, :h]                       # [n, j']
            # fourier_s even part / odd part, [k, j']
            fe = snm[:, ev, m] @ q[ev, :]
            fo = snm[:, od, m] @ q[od, :] if od.size else np.zeros_like(fe)
            g[:, :h, m] = fe + fo
            g[:, nlat - 1:h - 1:-1, m] = fe - fo
        g[:, :, 0] /= 2.0
        pfield[...] = 2.0 * nlon * np.real(np.fft.ifft(g, axis=-1))
        return pfield

    def trans_grid_to_fourier(self, pfield, fourier_g):
        """SSM:295-330: fourier_g = fft(pfield)/nλ along longitude."""
        fourier_g[...] = np.fft.fft(pfield, axis=-1) / self.nlon
        return fourier_g

    def trans_grid_to_spherical(self, pfield, snm):
        """SSM:221-293: FFT/nλ, even/odd folded Legendre quadrature over
        degrees m..N (SSM:278), degree num_spherical row zeroed (SSM:290).
        Lower-triangle entries (n < m) are left untouched, as in SSM:277-288."""
        self._check_levels(pfield.shape[0])
        N, ns = self.num_fourier, self.num_spherical
        nlat = self.nlat
        assert nlat % 2 == 0
        h = nlat // 2
        g = np.fft.fft(pfield, axis=-1) / self.nlon
        for m in range(N + 1):
            gt = g[:, :, m]                              # [k, j]
            e = gt[:, :h] + gt[:, nlat - 1:h - 1:-1]
            o = gt[:, :h] - gt[:, nlat - 1:h - 1:-1]
            for n in range(m, ns):
                src = e if (n - m) % 2 == 0 else o
                snm[:, n, m] = src @ self.qwg[m, n, :h] / 2.0
        snm[:, ns, :] = 0.0
        return snm

    # ------------------------------------------------------------ operators
    def divide_by_cos(self, grid):
        grid /= self.cos_theta[None, :, None]            # SSM:333-338
        return grid

    def multiply_by_cos(self, grid, out):
        out[...] = grid * self.cos_theta[None, :, None]  # SSM:341-346
        return out

    def compute_alpha_operator(self, fourier_a, fourier_b, isign, alpha):
        """SSM:424-459 over all latitudes, degrees m..num_spherical:
        alpha = 1/(2r) (i A . coef_a + isign B . coef_b)."""
        N, ns = self.num_fourier, self.num_spherical
        for m in range(N + 1):
            A = fourier_a[:, :, m]                       # [k, j]
            B = fourier_b[:, :, m]
            n = slice(m, ns + 1)
            alpha[:, n, m] = 1.0 / (2.0 * self.radius) * (
                (1j * A) @ self.coef_alp_a[m, n, :].T
                + isign * (B @ self.coef_alp_b[m, n, :].T))
        return alpha

    def vor_div_from_grid_uv(self, grid_u, grid_v, vor, div):
        """SSM:349-393."""
        self._check_levels(grid_u.shape[0])
        ns = self.num_spherical
        nlev = grid_u.shape[0]
        ucos = self.multiply_by_cos(grid_u, np.empty_like(grid_u))
        vcos = self.multiply_by_cos(grid_v, np.empty_like(grid_v))
        fu = self.trans_grid_to_fourier(ucos, np.empty((nlev, self.nlat, self.nlon), np.complex128))
        fv = self.trans_grid_to_fourier(vcos, np.empty((nlev, self.nlat, self.nlon), np.complex128))
        self.compute_alpha_operator(fv, fu, -1.0, vor)
        self.compute_alpha_operator(fu, fv, 1.0, div)
        vor[:, ns, :] = 0.0
        div[:, ns, :] = 0.0
        return vor, div

    def compute_ucos_vcos_from_vor_div(self, vor, div, ucos, vcos):
        """SSM:517-580 (all rows, the top degree included)."""
        ns = self.num_spherical
        uvc, uvm, uvp = self.coef_uvc.T, self.coef_uvm.T, self.coef_uvp.T   # [n, m]
        ucos[...] = uvc * (1j * div)
        ucos[:, 1:ns + 1, :] -= uvm[1:ns + 1, :] * vor[:, 0:ns, :]
        ucos[:, 0:ns, :] -= uvp[0:ns, :] * vor[:, 1:ns + 1, :]
        vcos[...] = uvc * (1j * vor)
        vcos[:, 1:ns + 1, :] += uvm[1:ns + 1, :] * div[:, 0:ns, :]
        vcos[:, 0:ns, :] += uvp[0:ns, :] * div[:, 1:ns + 1, :]
        return ucos, vcos

    def uv_grid_from_vor_div(self, vor, div, grid_u, grid_v):
        """SSM:462-482."""
        self._check_levels(vor.shape[0])
        ucos = np.empty_like(vor)
        vcos = np.empty_like(vor)
        self.compute_ucos_vcos_from_vor_div(vor, div, ucos, vcos)
        self.trans_spherical_to_grid(ucos, grid_u)
        self.trans_spherical_to_grid(vcos, grid_v)
        self.divide_by_cos(grid_u)
        self.divide_by_cos(grid_v)
        return grid_u, grid_v

    def apply_laplacian(self, s):
        s *= self.laplacian_eig.T[None, :, :]            # SSM:614-621
        return s

    def compute_gradient_cos(self, hs, dlam_hs, dth_hs):
        """SSM:655-692."""
        ns = self.num_spherical
        dlam_hs[...] = self.coef_dlam.T * (1j * hs)
        dth_hs[:, ns, :] = 0.0
        dth_hs[:, 0:ns, :] = self.coef_dthp.T[0:ns, :] * hs[:, 1:ns + 1, :]
        dth_hs[:, 1:ns + 1, :] += self.coef_dthm.T[1:ns + 1, :] * hs[:, 0:ns, :]
        return dlam_hs, dth_hs

    def compute_gradients(self, hs, grid_dlam, grid_dth):
        """SSM:721-741."""
        self._check_levels(hs.shape[0])
        a = np.empty_like(hs)
        b = np.empty_like(hs)
        self.compute_gradient_cos(hs, a, b)
        self.trans_spherical_to_grid(a, grid_dlam)
        self.trans_spherical_to_grid(b, grid_dth)
        self.divide_by_cos(grid_dlam)
        self.divide_by_cos(grid_dth)
        return grid_dlam, grid_dth

    def add_horizontal_advection(self, hs, grid_u, grid_v, grid_dh):
        """SSM:694-718: grid_δh -= u dλh + v dθh."""
        self._check_levels(hs.shape[0])
        gl = np.empty((hs.shape[0], self.nlat, self.nlon))
        gt = np.empty_like(gl)
        self.compute_gradients(hs, gl, gt)
        grid_dh -= (grid_u * gl + grid_v * gt)
        return grid_dh

    def area_weighted_global_mean(self, grid):
        """SSM:746-760: level 1 only, sum_ij f w_j / (2 nλ)."""
        return float(np.sum(grid[0].T @ self.wts) / (2.0 * self.nlon))
