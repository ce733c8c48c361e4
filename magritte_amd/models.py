"""Model data of the formal solver and seeded synthetic model generators.

``ModelData`` holds the arrays the reference's ``Model`` carries on the
formal-solver path (geometry, thermodynamics, chemistry, line data,
quadrature) and runs the host-side producers of libmgx
(``mgx_compute_*``) to derive the solver inputs: spectral grid, inverse
widths, LTE populations, emissivities/opacities, dshift_max and the
interpolation criterion.

The generators restate the conventions of the reference's model-building
helpers (magritte/setup.py): 1D neighbour lists (setup.py:45-98), Delaunay
neighbour lists (setup.py:84-92), HEALPix directions rotated so the first ray
points along +x (setup.py:101-122, 174-223), Gauss-Hermite quadrature with
weights normalised to one (setup.py:480-503) and LAMDA-style derived line data
(setup.py:759-767).  HEALPix is computed here (healpy is not available).
"""
import ctypes as C

import numpy as np

from . import _abi

# src/tools/constants.hpp
CC = 2.99792458e8
HH = 6.62607004e-34
KB = 1.38064852e-23
AMU = 1.66053904e-27
T_CMB = 2.72548
CC_SQUARED = 8.98755179e16  # the rounded literal of constants.hpp

# CGS constants for Einstein A of a rotor
_H_CGS = 6.62607004e-27
_C_CGS = 2.99792458e10
_DEBYE = 1.0e-18

SEED_BASE = 20261202


def seed_for(config, model_index=0):
    return SEED_BASE + 1000 * config + model_index


# ---------------------------------------------------------------------------
# line data


def linedata_two_level(freq, A, g=(1.0, 3.0), mass=28.0, num=1):
    """A two-level species with line frequency ``freq`` (setup.py:759-767 derived terms)."""
    energy = np.array([0.0, HH * freq])
    weight = np.array(g, dtype=np.float64)
    return _linedata(num, mass, energy, weight, [1], [0], [A])


def linedata_linear_rotor(B, dipole_debye, nlev, mass=28.0, num=1):
    """Linear rotor: E_J = h B J(J+1), g_J = 2J+1, J -> J-1 lines, A from the dipole."""
    J = np.arange(nlev, dtype=np.float64)
    energy = HH * B * J * (J + 1.0)
    weight = 2.0 * J + 1.0
    irad = list(range(1, nlev))
    jrad = list(range(0, nlev - 1))
    mu = dipole_debye * _DEBYE
    A = []
    for Ju in irad:
        nu = 2.0 * B * Ju
        A.append(64.0 * np.pi**4 * nu**3 * mu**2 / (3.0 * _H_CGS * _C_CGS**3) * Ju / (2.0 * Ju + 1.0))
    return _linedata(num, mass, energy, weight, irad, jrad, A)


def linedata_blended(freq, separations, A=1.0e-4, mass=28.0, num=1):
    """A ladder species whose lines overlap: level k+1 lies h*nu_k above level k,
    with nu_0 = freq and nu_k = freq * (1 + separations[k-1]) (relative
    separations, e.g. a few Doppler widths); transitions k+1 -> k with Einstein
    A (scalar or one per line) and weights 2k+1.  Lines in ascending frequency."""
    nus = [freq] + [freq * (1.0 + s) for s in separations]
    energy = np.concatenate([[0.0], np.cumsum([HH * nu for nu in nus])])
    nlev = len(energy)
    weight = 2.0 * np.arange(nlev, dtype=np.float64) + 1.0
    irad = list(range(1, nlev))
    jrad = list(range(0, nlev - 1))
    Av = [A] * len(irad) if np.isscalar(A) else list(A)
    return _linedata(num, mass, energy, weight, irad, jrad, Av)


def _linedata(num, mass, energy, weight, irad, jrad, A):
    irad = np.asarray(irad, dtype=np.uint32)
    jrad = np.asarray(jrad, dtype=np.uint32)
    A = np.asarray(A, dtype=np.float64)
    nrad = len(irad)
    freq = np.empty(nrad)
    Bs = np.empty(nrad)
    Ba = np.empty(nrad)
    for k in range(nrad):
        i, j = int(irad[k]), int(jrad[k])
        freq[k] = (energy[i] - energy[j]) / HH
        Bs[k] = A[k] * CC**2 / (2.0 * HH * freq[k] ** 3)
        Ba[k] = weight[i] / weight[j] * Bs[k]
    return dict(num=num, nlev=len(energy), nrad=nrad, inverse_mass=float(1.0 / mass),
                energy=np.ascontiguousarray(energy, dtype=np.float64),
                weight=np.ascontiguousarray(weight, dtype=np.float64),
                irad=irad, jrad=jrad, A=A, Bs=Bs, Ba=Ba, frequency=freq)


def colpar_constant(ld, num_col_partner, rate, tmp=None, orth_or_para="n"):
    """A collision partner with the same de-excitation coefficient ``rate`` [m^3/s]
    for every level pair (upper i > lower j) at every tabulated temperature, and
    excitation coefficients from detailed balance at each of them,
    Ce = Cd g_i/g_j exp(-(E_i - E_j)/(k T)) (the derivation of setup.py:722-728)."""
    tmp = np.geomspace(5.0, 2.0e4, 13) if tmp is None else np.asarray(tmp, dtype=np.float64)
    nlev = ld["nlev"]
    icol, jcol = [], []
    for i in range(1, nlev):
        for j in range(i):
            icol.append(i)
            jcol.append(j)
    return _colpar(ld, num_col_partner, orth_or_para, icol, jcol, tmp,
                   np.full((len(tmp), len(icol)), float(rate)))


def _colpar(ld, num_col_partner, orth_or_para, icol, jcol, tmp, Cd):
    icol = np.ascontiguousarray(icol, dtype=np.uint32)
    jcol = np.ascontiguousarray(jcol, dtype=np.uint32)
    tmp = np.ascontiguousarray(tmp, dtype=np.float64)
    Cd = np.ascontiguousarray(Cd, dtype=np.float64).reshape(len(tmp), len(icol))
    E, g = ld["energy"], ld["weight"]
    Ce = np.empty_like(Cd)
    for t in range(len(tmp)):
        for k in range(len(icol)):
            i, j = int(icol[k]), int(jcol[k])
            Ce[t, k] = Cd[t, k] * g[i] / g[j] * np.exp(-(E[i] - E[j]) / (KB * tmp[t]))
    return dict(num_col_partner=int(num_col_partner), orth_or_para=orth_or_para, icol=icol,
                jcol=jcol, tmp=tmp, Cd=Cd, Ce=Ce)


def quadrature(nquads):
    roots, weights = np.polynomial.hermite.hermgauss(nquads)
    return roots, weights / np.sum(weights)


# ---------------------------------------------------------------------------
# HEALPix (RING scheme) pixel centres


def healpix_pix2vec(nside):
    npix = 12 * nside * nside
    ncap = 2 * nside * (nside - 1)
    out = np.empty((npix, 3))
    for ip in range(npix):
        if ip < ncap:
            iring = int((1 + int(np.sqrt(1 + 2 * ip))) // 2)
            while 2 * iring * (iring - 1) > ip:
                iring -= 1
            while 2 * (iring + 1) * iring <= ip:
                iring += 1
            iphi = ip - 2 * iring * (iring - 1) + 1
            z = 1.0 - iring * iring / (3.0 * nside * nside)
            phi = (iphi - 0.5) * np.pi / (2.0 * iring)
        elif ip < npix - ncap:
            q = ip - ncap
            iring = q // (4 * nside) + nside
            iphi = q % (4 * nside) + 1
            fodd = 0.5 * (1 + ((iring + nside) & 1))
            z = (2 * nside - iring) * 2.0 / (3.0 * nside)
            phi = (iphi - fodd) * np.pi / (2.0 * nside)
        else:
            q = npix - ip
            iring = int((1 + int(np.sqrt(2 * q - 1))) // 2)
            while 2 * iring * (iring - 1) >= q:
                iring -= 1
            while 2 * (iring + 1) * iring < q:
                iring += 1
            iphi = 4 * iring + 1 - (q - 2 * iring * (iring - 1))
            z = -1.0 + iring * iring / (3.0 * nside * nside)
            phi = (iphi - 0.5) * np.pi / (2.0 * iring)
        sth = np.sqrt(max(0.0, (1.0 - z) * (1.0 + z)))
        out[ip] = (sth * np.cos(phi), sth * np.sin(phi), z)
    return out


def _rotation_matrix(n):
    """setup.py:101-122"""
    nx = np.array([1.0, 0.0, 0.0])
    ny = np.array([0.0, 1.0, 0.0])
    n1 = np.array(n, dtype=np.float64)
    n1 /= np.linalg.norm(n1)
    if np.linalg.norm(n1 - nx) < 1.0e-6:
        n2 = np.cross(ny, n1)
    else:
        n2 = np.cross(nx, n1)
    n2 /= np.linalg.norm(n2)
    n3 = np.cross(n2, n1)
    n3 /= np.linalg.norm(n3)
    return np.array([n1, n2, n3]).T


def random_rotations(n, seed):
    """n seeded rotation matrices, uniform over SO(3) (unit quaternions)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    q = rng.standard_normal((n, 4))
    q /= np.linalg.norm(q, axis=1)[:, None]
    w, x, y, z = q.T
    R = np.empty((n, 3, 3))
    R[:, 0, 0] = 1 - 2 * (y * y + z * z)
    R[:, 0, 1] = 2 * (x * y - z * w)
    R[:, 0, 2] = 2 * (x * z + y * w)
    R[:, 1, 0] = 2 * (x * y + z * w)
    R[:, 1, 1] = 1 - 2 * (x * x + z * z)
    R[:, 1, 2] = 2 * (y * z - x * w)
    R[:, 2, 0] = 2 * (x * z - y * w)
    R[:, 2, 1] = 2 * (y * z + x * w)
    R[:, 2, 2] = 1 - 2 * (x * x + y * y)
    return R


def rotated_directions(rotations, directions):
    """directions (nrays, 3) rotated by each of rotations (n, 3, 3): (n, nrays, 3),
    with -d mapped to exactly -(R d), so antipodes stay antipodes."""
    d = np.asarray(directions, dtype=np.float64)
    out = np.empty((len(rotations), len(d), 3))
    for i in range(3):
        out[:, :, i] = (rotations[:, i, 0][:, None] * d[None, :, 0]
                        + rotations[:, i, 1][:, None] * d[None, :, 1]
                        + rotations[:, i, 2][:, None] * d[None, :, 2])
    return out


def uniform_rays(nrays, first_ray=(1.0, 0.0, 0.0)):
    """HEALPix directions rotated so ray 0 points along first_ray (setup.py:174-223)."""
    nside = int(round(np.sqrt(nrays / 12)))
    assert 12 * nside * nside == nrays, "nrays must be 12*nside^2"
    d = healpix_pix2vec(nside)
    R1 = _rotation_matrix(d[0])
    R2 = _rotation_matrix(np.asarray(first_ray, dtype=np.float64))
    R = R2 @ np.linalg.inv(R1)
    return np.ascontiguousarray(d @ R.T)


# ---------------------------------------------------------------------------


class ModelData:
    """Arrays of one model on the formal-solver path (host numpy)."""

    def __init__(self):
        self.dimension = 3
        self.spherical_symmetry = 0
        self.adaptive_directions = 0
        self.linedata = []
        self.params = dict(max_width_fraction=0.35, min_opacity=1.0e-26,
                           min_line_opacity=1.0e-13, min_line_emissivity=1.0e-50,
                           max_interpolation_diff=1.4, max_distance_opacity_contribution=10.0)

    # -- geometry ---------------------------------------------------------
    def set_geometry(self, position, velocity, neighbors_list, boundary, directions, weights,
                     boundary_condition=None, boundary_temperature=None):
        P = len(position)
        self.npoints = P
        self.position = np.ascontiguousarray(position, dtype=np.float64)
        self.velocity = np.ascontiguousarray(velocity, dtype=np.float64)
        if isinstance(neighbors_list, tuple):  # CSR (n_neighbors, neighbors)
            n_nbs, nbs = neighbors_list
        else:
            n_nbs = [len(x) for x in neighbors_list]
            nbs = [v for x in neighbors_list for v in x]
        self.n_neighbors = np.ascontiguousarray(n_nbs, dtype=np.uint32)
        self.neighbors = np.ascontiguousarray(nbs, dtype=np.uint32)
        cum = np.zeros(P, dtype=np.uint32)
        cum[1:] = np.cumsum(self.n_neighbors[:-1], dtype=np.uint64).astype(np.uint32)
        self.cum_n_neighbors = cum
        self.boundary2point = np.ascontiguousarray(boundary, dtype=np.uint32)
        self.nboundary = len(self.boundary2point)
        p2b = np.full(P, P, dtype=np.uint32)
        p2b[self.boundary2point] = np.arange(self.nboundary, dtype=np.uint32)
        self.point2boundary = p2b
        nb = self.nboundary
        self.boundary_condition = np.ascontiguousarray(
            np.full(nb, _abi.BC_CMB) if boundary_condition is None else boundary_condition,
            dtype=np.int32)
        self.boundary_temperature = np.ascontiguousarray(
            np.full(nb, T_CMB) if boundary_temperature is None else boundary_temperature,
            dtype=np.float64)
        self.ray_direction = np.ascontiguousarray(directions, dtype=np.float64)
        self.ray_weight = np.ascontiguousarray(weights, dtype=np.float64)
        self.nrays = len(self.ray_weight) if not self.adaptive_directions else self.ray_weight.size // P
        self.hnrays = self.nrays // 2
        # antipods (rays.cpp:52-61): last r2 with |d1 + d2|^2 < 1e-9
        d = self.ray_direction.reshape(-1, 3)[: self.nrays]
        anti = np.zeros(self.nrays, dtype=np.uint32)
        for r1 in range(self.nrays):
            for r2 in range(self.nrays):
                s = d[r1] + d[r2]
                if s[0] * s[0] + s[1] * s[1] + s[2] * s[2] < 1.0e-9:
                    anti[r1] = r2
        self.antipod = anti
        # Points::center (points.cpp:29-48), including its max initialisation
        lo = np.full(3, np.finfo(np.float64).max)
        hi = np.full(3, np.finfo(np.float64).tiny)
        lo = np.minimum(lo, self.position.min(axis=0))
        hi = np.maximum(hi, self.position.max(axis=0))
        self.center = (lo + hi) / 2

    def set_adaptive_directions(self, directions, weights):
        """Per-point ray directions (Rays::use_adaptive_directions, rays.cpp:14-17,
        37-50): directions (P, nrays, 3) and weights (P, nrays), stored point-major
        as the reference reads them ((P*nrays, 3) and (P*nrays)).  The antipod
        table comes from the first point's directions (rays.cpp:52-61: it is
        assumed the same for every point)."""
        P = self.npoints
        d = np.asarray(directions, dtype=np.float64)
        w = np.asarray(weights, dtype=np.float64)
        assert d.shape[0] == P and d.shape[2] == 3 and w.shape == d.shape[:2]
        self.adaptive_directions = 1
        self.ray_direction = np.ascontiguousarray(d.reshape(-1, 3))
        self.ray_weight = np.ascontiguousarray(w.reshape(-1))
        self.nrays = d.shape[1]
        self.hnrays = self.nrays // 2
        anti = np.zeros(self.nrays, dtype=np.uint32)
        d0 = d[0]
        for r1 in range(self.nrays):
            for r2 in range(self.nrays):
                t = d0[r1] + d0[r2]
                if t[0] * t[0] + t[1] * t[1] + t[2] * t[2] < 1.0e-9:
                    anti[r1] = r2
        self.antipod = anti
        return self

    # -- physics ----------------------------------------------------------
    def set_thermo(self, temperature, vturb, abundance):
        self.temperature = np.ascontiguousarray(temperature, dtype=np.float64)
        self.vturb2 = np.ascontiguousarray((np.asarray(vturb, dtype=np.float64) / CC) ** 2)
        self.abundance = np.ascontiguousarray(abundance, dtype=np.float64)
        self.nspec = self.abundance.shape[1]

    def set_lines(self, linedata, nquads):
        self.linedata = list(linedata)
        self.nlspecs = len(self.linedata)
        self.nquads = nquads
        roots, weights = quadrature(nquads)
        self.quad_roots = np.ascontiguousarray(np.tile(roots, (self.nlspecs, 1)))
        self.quad_weights = np.ascontiguousarray(np.tile(weights, (self.nlspecs, 1)))

    def _ld_array(self):
        arr = (_abi.MgxLinedata * self.nlspecs)()
        keep = []
        for s, d in enumerate(self.linedata):
            a = arr[s]
            a.num, a.nlev, a.nrad = d["num"], d["nlev"], d["nrad"]
            a.inverse_mass = d["inverse_mass"]
            for f in ("energy", "weight", "A", "Bs", "Ba", "frequency"):
                x = np.ascontiguousarray(d[f], dtype=np.float64)
                keep.append(x)
                setattr(a, f, x.ctypes.data_as(C.POINTER(C.c_double)))
            for f in ("irad", "jrad"):
                x = np.ascontiguousarray(d[f], dtype=np.uint32)
                keep.append(x)
                setattr(a, f, x.ctypes.data_as(C.POINTER(C.c_uint32)))
        return arr, keep

    def _colpar_arrays(self):
        """(ncolpar (nlspecs,) uint32, MgxColldata array in species order, keep-alive list)."""
        ncol = np.ascontiguousarray([len(d.get("colpar", [])) for d in self.linedata],
                                    dtype=np.uint32)
        allc = [c for d in self.linedata for c in d.get("colpar", [])]
        arr = (_abi.MgxColldata * max(1, len(allc)))()
        keep = [ncol]
        for a, c in zip(arr, allc):
            a.num_col_partner = c["num_col_partner"]
            a.orth_or_para_H2 = ord(c["orth_or_para"])
            a.ntmp, a.ncol = len(c["tmp"]), len(c["icol"])
            for f in ("icol", "jcol"):
                x = np.ascontiguousarray(c[f], dtype=np.uint32)
                keep.append(x)
                setattr(a, f, x.ctypes.data_as(C.POINTER(C.c_uint32)))
            for f in ("tmp", "Cd", "Ce"):
                x = np.ascontiguousarray(c[f], dtype=np.float64)
                keep.append(x)
                setattr(a, f, x.ctypes.data_as(C.POINTER(C.c_double)))
        return ncol, arr, keep

    def finalize(self):
        """Run the host producers (Model::compute_* on the reference's path)."""
        L = _abi.lib()
        P, Q = self.npoints, self.nquads
        ld, keep = self._ld_array()
        line, lsp, lk, lA = [], [], [], []
        for s, d in enumerate(self.linedata):
            for k in range(d["nrad"]):
                line.append(d["frequency"][k])
                lsp.append(s)
                lk.append(k)
                lA.append(d["A"][k])
        self.nlines = len(line)
        self.line = np.ascontiguousarray(line, dtype=np.float64)
        self.line_species = np.ascontiguousarray(lsp, dtype=np.uint32)
        self.line_k = np.ascontiguousarray(lk, dtype=np.uint32)
        self.line_A = np.ascontiguousarray(lA, dtype=np.float64)
        self.species_inverse_mass = np.ascontiguousarray(
            [d["inverse_mass"] for d in self.linedata], dtype=np.float64)
        self.max_inverse_mass = float(max(0.0, *[d["inverse_mass"] for d in self.linedata]))
        self.sorted_line = np.empty(self.nlines)
        self.sorted_line_map = np.empty(self.nlines, dtype=np.uint32)
        _abi.check(L.mgx_sort_lines(self.nlines, _abi.ptr(self.line), _abi.ptr(self.sorted_line),
                                    _abi.ptr(self.sorted_line_map)))
        self.compute_spectral_discretisation()
        self.inverse_width = np.empty((P, self.nlines))
        _abi.check(L.mgx_compute_inverse_widths(P, self.nlspecs, ld, _abi.ptr(self.temperature),
                                                _abi.ptr(self.vturb2), _abi.ptr(self.inverse_width)))
        self.compute_lte()
        self.dshift_max = np.empty(P)
        _abi.check(L.mgx_compute_dshift_max(P, self.nlspecs, ld, _abi.ptr(self.temperature),
                                            _abi.ptr(self.vturb2),
                                            self.params["max_width_fraction"],
                                            _abi.ptr(self.dshift_max)))
        del keep
        return self

    def compute_spectral_discretisation(self):
        L = _abi.lib()
        ld, keep = self._ld_array()
        P, Q = self.npoints, self.nquads
        F = self.nlines * Q
        self.nfreqs = F
        self.nu = np.empty((P, F))
        self.f_line = np.empty(F, dtype=np.uint32)
        self.f_species = np.empty(F, dtype=np.uint32)
        self.f_k = np.empty(F, dtype=np.uint32)
        self.f_z = np.empty(F, dtype=np.uint32)
        self.f_in_line = np.empty(F, dtype=np.uint8)
        _abi.check(L.mgx_compute_spectral_discretisation(
            P, self.nlspecs, ld, Q, _abi.ptr(self.quad_roots), _abi.ptr(self.temperature),
            _abi.ptr(self.vturb2), _abi.ptr(self.nu), _abi.ptr(self.f_line),
            _abi.ptr(self.f_species), _abi.ptr(self.f_k), _abi.ptr(self.f_z),
            _abi.ptr(self.f_in_line)))
        del keep

    def compute_lte(self):
        L = _abi.lib()
        ld, keep = self._ld_array()
        P = self.npoints
        npop = sum(d["nlev"] for d in self.linedata) * P
        self.population = np.empty(npop)
        self.population_tot = np.empty((self.nlspecs, P))
        self.emissivity = np.empty((P, self.nlines))
        self.opacity = np.empty((P, self.nlines))
        _abi.check(L.mgx_compute_lte(P, self.nspec, _abi.ptr(self.abundance),
                                     _abi.ptr(self.temperature), self.nlspecs, ld,
                                     self.params["min_line_emissivity"],
                                     self.params["min_line_opacity"], _abi.ptr(self.population),
                                     _abi.ptr(self.population_tot), _abi.ptr(self.emissivity),
                                     _abi.ptr(self.opacity)))
        self.interp_crit = np.empty((P, self.nlines))
        _abi.check(L.mgx_compute_interp_crit(P, self.nlspecs, ld, _abi.ptr(self.opacity),
                                             _abi.ptr(self.population_tot),
                                             _abi.ptr(self.interp_crit)))
        del keep

    # -- views ------------------------------------------------------------
    ARRAYS = ("position", "velocity", "n_neighbors", "cum_n_neighbors", "neighbors",
              "point2boundary", "boundary2point", "boundary_condition", "boundary_temperature",
              "ray_direction", "ray_weight", "antipod", "temperature", "vturb2", "line",
              "sorted_line", "sorted_line_map", "emissivity", "opacity", "inverse_width",
              "interp_crit", "dshift_max", "line_species", "line_k", "line_A",
              "species_inverse_mass", "quad_weights", "nu", "f_line", "f_species", "f_k",
              "f_z", "f_in_line")

    def options(self, approx=_abi.APPROX_CLOSE_LINES, precision=_abi.PREC_FP64):
        o = _abi.MgxOptions()
        o.approx = approx
        o.n_off_diag = 0
        o.precision = precision
        o.min_opacity = self.params["min_opacity"]
        o.max_distance_opacity_contribution = self.params["max_distance_opacity_contribution"]
        o.max_interpolation_diff = self.params["max_interpolation_diff"]
        return o

    FREQ_ARRAYS = ("nu", "f_line", "f_species", "f_k", "f_z", "f_in_line")

    def image_frequencies(self, nu_min, nu_max, n):
        """Model::compute_spectral_discretisation(nu_min, nu_max, n) (model.cpp:257-298):
        the arrays of an image frequency grid (the line grid of the model is kept)."""
        P = self.npoints
        d = dict(nu=np.empty((P, n)), f_line=np.empty(n, dtype=np.uint32),
                 f_species=np.empty(n, dtype=np.uint32), f_k=np.empty(n, dtype=np.uint32),
                 f_z=np.empty(n, dtype=np.uint32), f_in_line=np.empty(n, dtype=np.uint8))
        _abi.check(_abi.lib().mgx_compute_image_discretisation(
            P, float(nu_min), float(nu_max), n, *[_abi.ptr(d[k]) for k in self.FREQ_ARRAYS]))
        return d

    def view(self, arrays=None, nfreqs=None):
        """mgx_model whose pointers refer to ``arrays`` (default: the numpy arrays);
        entries missing from ``arrays`` are taken from the model."""
        base = {k: getattr(self, k) for k in self.ARRAYS}
        if arrays is not None:
            base.update(arrays)
        arrays = base
        m = _abi.MgxModel()
        for f in ("npoints", "nrays", "hnrays", "nboundary", "nlines", "nlspecs", "nquads",
                  "nfreqs", "dimension", "spherical_symmetry", "adaptive_directions"):
            setattr(m, f, int(getattr(self, f)))
        if nfreqs is not None:
            m.nfreqs = int(nfreqs)
        for k in self.ARRAYS:
            ftype = dict(m._fields_)[k]
            setattr(m, k, C.cast(C.c_void_p(_abi.ptr(arrays[k])), ftype))
        m.center[:] = [float(x) for x in self.center]
        m.max_inverse_mass = self.max_inverse_mass
        return m

    def to_core(self):
        """magritte_amd.core.Model (the C++ Model with the reference's object model) built
        from this model as the reference's scripts build theirs: parameters.set_*, .set()
        on the geometry and thermodynamics vectors, the abundance table, the line data and
        quadrature of every species; then its line grid and widths are computed and the
        line emissivity, opacity and populations copied."""
        from . import core

        cm = core.Model()
        p = cm.parameters
        P = self.npoints
        p.set_dimension(int(self.dimension))
        p.set_npoints(P)
        p.set_nrays(int(self.nrays))
        p.set_nspecs(int(self.nspec))
        p.set_nlspecs(int(self.nlspecs))
        p.set_nquads(int(self.nquads))
        p.set_spherical_symmetry(bool(self.spherical_symmetry))
        for k in ("min_opacity", "max_distance_opacity_contribution", "max_interpolation_diff",
                  "min_line_emissivity", "min_line_opacity", "max_width_fraction"):
            setattr(p, k, float(self.params[k]))
        g = cm.geometry
        g.points.position.set(self.position.reshape(P, 3))
        g.points.velocity.set(self.velocity.reshape(P, 3))
        g.points.n_neighbors.set(self.n_neighbors)
        g.points.neighbors.set(self.neighbors)
        g.rays.direction.set(self.ray_direction.reshape(-1, 3))
        g.rays.weight.set(self.ray_weight)
        g.rays.antipod.set(self.antipod)
        g.boundary.boundary2point.set(self.boundary2point)
        g.boundary.boundary_temperature.set(self.boundary_temperature)
        for b, c in enumerate(self.boundary_condition):
            g.boundary.set_boundary_condition(b, core.BoundaryCondition(int(c)))
        cm.thermodynamics.temperature.gas.set(self.temperature)
        cm.thermodynamics.turbulence.vturb2.set(self.vturb2)
        cm.chemistry.species.abundance = np.ascontiguousarray(self.abundance)
        cm.lines.resize_LineProducingSpecies(self.nlspecs)
        off = 0
        for s, d in enumerate(self.linedata):
            ls = cm.lines.lineProducingSpecies[s]
            ld = ls.linedata
            ld.num, ld.nlev, ld.nrad = int(d["num"]), int(d["nlev"]), int(d["nrad"])
            ld.inverse_mass = float(d["inverse_mass"])
            for f in ("energy", "weight", "A", "Bs", "Ba", "frequency"):
                setattr(ld, f, [float(x) for x in d[f]])
            for f in ("irad", "jrad"):
                setattr(ld, f, [int(x) for x in d[f]])
            for c in d.get("colpar", []):
                cp = core.CollisionPartner()
                cp.num_col_partner = int(c["num_col_partner"])
                cp.orth_or_para_H2 = c["orth_or_para"]
                cp.icol, cp.jcol = [int(x) for x in c["icol"]], [int(x) for x in c["jcol"]]
                cp.tmp = [float(x) for x in c["tmp"]]
                cp.Cd = [float(x) for x in np.ravel(c["Cd"])]
                cp.Ce = [float(x) for x in np.ravel(c["Ce"])]
                ld.colpar.append(cp)
            ls.quadrature.roots.set(self.quad_roots[s])
            ls.quadrature.weights.set(self.quad_weights[s])
            n = self.npoints * int(d["nlev"])
            if getattr(self, "population", None) is not None:
                ls.population = self.population[off:off + n]
                ls.population_tot = self.population_tot[s]
            off += n
        cm.compute_spectral_discretisation()
        cm.compute_inverse_line_widths()
        cm.lines.emissivity.set(self.emissivity)
        cm.lines.opacity.set(self.opacity)
        return cm

    @classmethod
    def from_core(cls, cm):
        """The device view's arrays of a magritte_amd.core.Model (as it packs them for a
        solve) in a ModelData, e.g. to run the oracle on a model built through the
        reference's object model."""
        m = cls()
        u32, i32, f64, u8 = np.uint32, np.int32, np.float64, np.uint8
        kinds = {"boundary_condition": i32, "f_in_line": u8}
        for k in cls.ARRAYS:
            raw = bytes(cm.get_array(k))
            dt = kinds.get(k, f64 if k in cls._F64 else u32)
            setattr(m, k, np.frombuffer(raw, dtype=dt).copy())
        p = cm.parameters
        m.npoints, m.nrays, m.hnrays, m.nboundary = p.npoints(), p.nrays(), p.hnrays(), p.nboundary()
        m.nlines, m.nlspecs, m.nquads, m.nfreqs = p.nlines(), p.nlspecs(), p.nquads(), p.nfreqs()
        m.dimension = p.dimension()
        m.spherical_symmetry = int(p.spherical_symmetry())
        m.adaptive_directions = int(cm.geometry.rays.use_adaptive_directions)
        m.center = np.array(cm.geometry.points.center)
        m.max_inverse_mass = float(max(0.0, *m.species_inverse_mass))
        P = m.npoints
        m.position = m.position.reshape(P, 3)
        m.velocity = m.velocity.reshape(P, 3)
        m.ray_direction = m.ray_direction.reshape(-1, 3)
        m.nu = m.nu.reshape(P, -1)
        for k in ("emissivity", "opacity", "inverse_width", "interp_crit"):
            setattr(m, k, getattr(m, k).reshape(P, m.nlines))
        m.quad_weights = m.quad_weights.reshape(m.nlspecs, m.nquads)
        m.params.update(max_interpolation_diff=p.max_interpolation_diff,
                        min_opacity=p.min_opacity,
                        max_distance_opacity_contribution=p.max_distance_opacity_contribution,
                        max_width_fraction=p.max_width_fraction)
        return m

    _F64 = ("position", "velocity", "boundary_temperature", "ray_direction", "ray_weight",
            "temperature", "vturb2", "line", "sorted_line", "emissivity", "opacity",
            "inverse_width", "interp_crit", "dshift_max", "line_A", "species_inverse_mass",
            "quad_weights", "nu")

    def to_device(self, device="cuda"):
        """Copy every solver input to the GPU (torch tensors); returns (view, tensors)."""
        import torch
        t = {k: torch.from_numpy(np.ascontiguousarray(getattr(self, k))).to(device)
             for k in self.ARRAYS}
        return self.view(t), t


# ---------------------------------------------------------------------------
# synthetic workloads


def model_line_1d(npoints=100, length=1.0e13, nquads=31, vmax=1.0e4, density=1.0e12,
                  temperature=20.0, vturb=150.0, freq=115.27e9, A=7.2e-8):
    """1D line of points with neighbours i-1, i+1 and rays -x, +x; one two-level species."""
    P = npoints
    x = np.arange(P, dtype=np.float64) * (length / (P - 1))
    pos = np.zeros((P, 3))
    pos[:, 0] = x
    vel = np.zeros((P, 3))
    vel[:, 0] = (vmax * np.arange(P) / (P - 1)) / CC
    nbs = [[1]] + [[p - 1, p + 1] for p in range(1, P - 1)] + [[P - 2]]
    m = ModelData()
    m.dimension = 1
    m.set_geometry(pos, vel, nbs, [0, P - 1], [[-1.0, 0.0, 0.0], [1.0, 0.0, 0.0]], [0.5, 0.5])
    abund = np.zeros((P, 2))
    abund[:, 0] = density
    abund[:, 1] = 1.0e-4 * density
    m.set_thermo(np.full(P, temperature), np.full(P, vturb), abund)
    ld = linedata_two_level(freq, A)
    ld["colpar"] = [colpar_constant(ld, 0, 1.0e-16)]
    m.set_lines([ld], nquads)
    return m.finalize()


def jittered_grid(n_edge, half_size, rng, jitter=0.3):
    g = np.linspace(-half_size, half_size, n_edge)
    h = g[1] - g[0]
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    pos = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)
    pos = pos + rng.uniform(-jitter, jitter, size=pos.shape) * h
    I, Jg, K = np.meshgrid(np.arange(n_edge), np.arange(n_edge), np.arange(n_edge), indexing="ij")
    idx = np.stack([I.ravel(), Jg.ravel(), K.ravel()], axis=1)
    outer = np.any((idx == 0) | (idx == n_edge - 1), axis=1)
    return np.ascontiguousarray(pos), np.nonzero(outer)[0].astype(np.uint32)


def delaunay_neighbors(pos):
    from scipy.spatial import Delaunay

    indptr, nbs = Delaunay(pos).vertex_neighbor_vertices
    n_nbs = np.diff(indptr)
    return (n_nbs.astype(np.uint32), nbs.astype(np.uint32))


def grf_clumps(n_edge, sigma=0.5, corr_length=1.0 / 20.0, seed=None):
    """Log-normal clump factors exp(sigma g - sigma^2/2) on the n_edge^3 grid (the
    point order of jittered_grid), g a unit-variance Gaussian random field with a
    Gaussian correlation of length corr_length (fraction of the box): seeded
    white noise filtered in Fourier space."""
    rng = np.random.Generator(np.random.PCG64(seed_for(4) if seed is None else seed))
    w = rng.standard_normal((n_edge, n_edge, n_edge))
    k = 2.0 * np.pi * np.fft.fftfreq(n_edge)  # per grid spacing
    ell = corr_length * n_edge                # correlation length in grid spacings
    kx, ky, kz = np.meshgrid(k, k, k, indexing="ij")
    filt = np.exp(-0.25 * (kx**2 + ky**2 + kz**2) * ell**2)
    g = np.real(np.fft.ifftn(np.fft.fftn(w) * filt))
    g = (g - g.mean()) / g.std()
    return np.exp(sigma * g - 0.5 * sigma**2).ravel()


def model_wind(n_edge=32, half_size=1.0e14, nrays=12, nquads=31, seed=None, nlev=10,
               B=57.6e9, dipole=0.11, v_out=15.0e3, v_rot=0.0, vturb=150.0, clumps=None,
               temp_scale=None):
    """Jittered-grid spherical wind: n = 1e14 (r/1e12 m)^-2, T = 2000 (r/1e12 m)^-0.7,
    radial outflow, linear rotor; boundary = the outer grid layer (CMB)."""
    rng = np.random.Generator(np.random.PCG64(seed_for(1) if seed is None else seed))
    pos, bdy = jittered_grid(n_edge, half_size, rng)
    r = np.sqrt((pos**2).sum(axis=1))
    rhat = pos / r[:, None]
    vel = v_out * rhat
    if v_rot:
        vel = vel + v_rot * np.stack([-rhat[:, 1], rhat[:, 0], np.zeros_like(r)], axis=1)
    vel = vel / CC
    dens = 1.0e14 * (r / 1.0e12) ** -2
    if clumps is not None:
        dens = dens * clumps
    temp = 2000.0 * (r / 1.0e12) ** -0.7
    if temp_scale is not None:
        temp = temp * temp_scale
    m = ModelData()
    m.set_geometry(pos, vel, delaunay_neighbors(pos), bdy, uniform_rays(nrays),
                   np.full(nrays, 1.0 / nrays))
    abund = np.stack([dens, 1.0e-4 * dens], axis=1)
    m.set_thermo(temp, np.full(len(r), vturb), abund)
    ld = linedata_linear_rotor(B, dipole, nlev)
    ld["colpar"] = [colpar_constant(ld, 0, 1.0e-16)]
    m.set_lines([ld], nquads)
    return m.finalize()


def batch_wind_specs(n_models=64, p_min=2.0e3, p_max=1.3e5, nrays_choices=(12, 48, 192),
                     seed=None):
    """Sizes of a batch of jittered-grid winds: point counts log-uniform in
    [p_min, p_max] (grid edge round(P^(1/3))) and ray counts drawn from
    nrays_choices, seeded.  Returns a list of (n_edge, nrays, seed)."""
    rng = np.random.Generator(np.random.PCG64(seed_for(2) if seed is None else seed))
    logp = rng.uniform(np.log(p_min), np.log(p_max), n_models)
    nr = rng.choice(np.asarray(nrays_choices), n_models)
    return [(int(round(np.exp(lp) ** (1.0 / 3.0))), int(r), seed_for(2, i + 1))
            for i, (lp, r) in enumerate(zip(logp, nr))]


def _wind_from_spec(spec):
    e, r, sd = spec
    return model_wind(n_edge=e, nrays=r, seed=sd)


def batch_winds(specs, workers=1):
    """The wind models of batch_wind_specs (same rotor as model_wind), in order.
    workers > 1 builds them in that many forked processes (the Delaunay
    triangulation is single-threaded); fork before the process uses the GPU."""
    if workers <= 1 or len(specs) < 2:
        return [_wind_from_spec(sp) for sp in specs]
    import multiprocessing as mp

    # largest first, so the pool's tail is short
    order = sorted(range(len(specs)), key=lambda i: -specs[i][0])
    with mp.get_context("fork").Pool(min(workers, len(specs))) as pool:
        built = pool.map(_wind_from_spec, [specs[i] for i in order], chunksize=1)
    out = [None] * len(specs)
    for i, m in zip(order, built):
        out[i] = m
    return out


# LAMDA collision-partner codes -> (name, ortho/para flag)
LAMDA_PARTNERS = {1: ("H2", "n"), 2: ("H2", "p"), 3: ("H2", "o"), 4: ("e", "n"), 5: ("H", "n"),
                  6: ("He", "n"), 7: ("H+", "n")}


def read_lamda(path, num=1, species_index=None):
    """A LAMDA-format line data file as set_linedata_from_LAMDA_file stores it
    (setup.py:632-779): energies in J, levels 0-based, derived Bs, Ba, frequency;
    collision partners ("colpar") with Cd converted from cm^3/s to m^3/s and Ce
    from detailed balance at each tabulated temperature (setup.py:710-737).
    species_index maps partner names (LAMDA_PARTNERS) to abundance columns,
    default {"H2": 1}; an unknown partner gets column 0, the reference's dummy
    density (setup.py:568-572)."""
    species_index = {"H2": 1} if species_index is None else species_index
    with open(path) as fh:
        lines = [l.strip() for l in fh if l.strip()]
    data = [l for l in lines if not l.startswith("!")]
    mass = float(data[1])
    nlev = int(data[2])
    energy, weight = [], []
    for i in range(nlev):
        parts = data[3 + i].split()
        energy.append(float(parts[1]) * 1.0e2 * HH * CC)
        weight.append(float(parts[2]))
    nrad = int(data[3 + nlev])
    irad, jrad, A = [], [], []
    for k in range(nrad):
        parts = data[4 + nlev + k].split()
        irad.append(int(parts[1]) - 1)
        jrad.append(int(parts[2]) - 1)
        A.append(float(parts[3]))
    ld = _linedata(num, mass, np.array(energy), np.array(weight), irad, jrad, A)
    at = 4 + nlev + nrad
    colpar = []
    if at < len(data):
        ncolpar = int(data[at].split()[0])
        at += 1
        for _ in range(ncolpar):
            code = int(data[at].split()[0])
            ncol = int(data[at + 1].split()[0])
            ntmp = int(data[at + 2].split()[0])
            at += 3
            tmp = []
            while len(tmp) < ntmp:
                tmp.extend(float(t) for t in data[at].split())
                at += 1
            icol, jcol, Cd = [], [], np.empty((ntmp, ncol))
            for k in range(ncol):
                parts = data[at + k].split()
                icol.append(int(parts[1]) - 1)
                jcol.append(int(parts[2]) - 1)
                Cd[:, k] = [float(x) * 1.0e-6 for x in parts[3:3 + ntmp]]
            at += ncol
            name, op = LAMDA_PARTNERS.get(code, ("?", "n"))
            colpar.append(_colpar(ld, species_index.get(name, 0), op, icol, jcol, tmp, Cd))
    ld["colpar"] = colpar
    return ld


def model_constant_slab(npoints=50, dx=1.0e12, n_species=1.0e3, n_h2=1.0e12, temperature=4.5,
                        turb=0.0, nquads=1, linedata=None):
    """The reference's all-constant single-ray benchmark geometry
    (tests/benchmarks/analytic/all_constant_single_ray.py:18-70)."""
    P = npoints
    pos = np.zeros((P, 3))
    pos[:, 0] = np.arange(P) * dx
    vel = np.zeros((P, 3))
    nbs = [[1]] + [[p - 1, p + 1] for p in range(1, P - 1)] + [[P - 2]]
    m = ModelData()
    m.dimension = 1
    m.set_geometry(pos, vel, nbs, [0, P - 1], [[-1.0, 0.0, 0.0], [1.0, 0.0, 0.0]], [0.5, 0.5])
    abund = np.zeros((P, 3))
    abund[:, 0] = n_species
    abund[:, 1] = n_h2
    m.set_thermo(np.full(P, temperature), np.full(P, turb), abund)
    ld = linedata if linedata is not None else linedata_two_level(179.87547e9, 1.0e-4, mass=1.0, num=0)
    ld = dict(ld, num=0)
    m.set_lines([ld], nquads)
    return m.finalize()


def rays_spherical_symmetry(nrays):
    """Uniform ray set of a 1D spherically symmetric model (setup.py:226-339, uniform=True):
    hnrays directions from +x to +y in the xy-plane, weights from the covered solid angle."""
    nextra = nrays // 2 - 1
    Rx, Ry, Rz = [1.0], [0.0], [0.0]
    angle_max = np.arctan(Ry[-1] / Rx[-1])
    angle_extra = (0.5 * np.pi - angle_max) / nextra
    for k in range(1, nextra):
        Rx.append(np.cos(angle_max + k * angle_extra))
        Ry.append(np.sin(angle_max + k * angle_extra))
        Rz.append(0.0)
    Rx.append(0.0)
    Ry.append(1.0)
    Rz.append(0.0)
    n = 2 * len(Rx)
    Wt = []
    for i in range(n // 2):
        if i == 0:
            ux, uy = 0.5 * (Rx[i] + Rx[i + 1]), 0.5 * (Ry[i] + Ry[i + 1])
            lx, ly = Rx[0], Ry[0]
        elif i == n // 2 - 1:
            ux, uy = Rx[-1], Ry[-1]
            lx, ly = 0.5 * (Rx[i] + Rx[i - 1]), 0.5 * (Ry[i] + Ry[i - 1])
        else:
            ux, uy = 0.5 * (Rx[i] + Rx[i + 1]), 0.5 * (Ry[i] + Ry[i + 1])
            lx, ly = 0.5 * (Rx[i] + Rx[i - 1]), 0.5 * (Ry[i] + Ry[i - 1])
        Wt.append(lx / np.sqrt(lx**2 + ly**2) - ux / np.sqrt(ux**2 + uy**2))
    for i in range(n // 2):
        Rx.append(-Rx[i])
        Ry.append(-Ry[i])
        Rz.append(-Rz[i])
        Wt.append(Wt[i])
    d = np.array((Rx, Ry, Rz)).T
    w = np.array(Wt)
    return np.ascontiguousarray(d), w / np.sum(w)


def model_vanzadelhoff_1d(lamda_path, a_or_b="a", npoints=250, nrays=200, nquads=11):
    """The reference's van Zadelhoff 1 non-LTE benchmark in 1D spherical symmetry
    (tests/benchmarks/numeric/vanZadelhoff_1_1D.py:17-84): a two-level test
    species (LAMDA file) in a static r^-2 H2 envelope from 1e13 to 7.8e16 m,
    n_H2 = 2e13 (1e13 m / r)^2 m^-3, abundance 1e-8 ('a') or 1e-6 ('b'),
    T = 20 K, turbulence 150 m/s, CMB boundary, neighbours i-1, i+1."""
    r_in, r_out, nH2_in = 1.0e13, 7.8e16, 2.0e13
    X_mol = {"a": 1.0e-8, "b": 1.0e-6}[a_or_b]
    P = npoints
    rs = np.logspace(np.log10(r_in), np.log10(r_out), P, endpoint=True)
    pos = np.zeros((P, 3))
    pos[:, 0] = rs
    nbs = [[1]] + [[p - 1, p + 1] for p in range(1, P - 1)] + [[P - 2]]
    m = ModelData()
    m.dimension = 1
    m.spherical_symmetry = 1
    d, w = rays_spherical_symmetry(nrays)
    m.set_geometry(pos, np.zeros((P, 3)), nbs, [0, P - 1], d, w)
    nH2 = nH2_in * np.power(r_in / rs, 2.0)
    abund = np.stack([X_mol * nH2, nH2, np.zeros(P)], axis=1)
    m.set_thermo(np.full(P, 20.0), np.full(P, 150.0), abund)
    ld = read_lamda(lamda_path, num=0, species_index={"H2": 1})
    m.set_lines([ld], nquads)
    return m.finalize()


def model_velocity_gradient_1d(linedata, npoints=100, nrays=100, nquads=100, dx=1.0e4,
                               dv=2.5e2, n_species=1.0e8, n_h2=1.0e12, temperature=45.0,
                               turb=0.0, max_width_fraction=0.35):
    """The reference's constant-velocity-gradient 1D spherical benchmark
    (tests/benchmarks/analytic/constant_velocity_gradient_1D.py:18-77)."""
    P = npoints
    pos = np.zeros((P, 3))
    pos[:, 0] = (np.arange(P) + 1) * dx
    vel = np.zeros((P, 3))
    vel[:, 0] = (np.arange(P) + 1) * (dv / CC)
    nbs = [[1]] + [[p - 1, p + 1] for p in range(1, P - 1)] + [[P - 2]]
    m = ModelData()
    m.dimension = 1
    m.spherical_symmetry = 1
    m.params["max_width_fraction"] = max_width_fraction
    d, w = rays_spherical_symmetry(nrays)
    m.set_geometry(pos, vel, nbs, [0, P - 1], d, w)
    abund = np.zeros((P, 3))
    abund[:, 0] = n_species
    abund[:, 1] = n_h2
    m.set_thermo(np.full(P, temperature), np.full(P, turb), abund)
    m.set_lines([dict(linedata, num=0)], nquads)
    return m.finalize()
