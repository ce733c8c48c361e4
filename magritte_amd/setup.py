"""Model construction on the reference's object model (magritte_amd.core.Model),
with the names and arguments of the reference's setup helpers
(magritte/setup.py), so that scripts written against the reference build their
models unchanged with ``import magritte_amd.setup as setup``:

  set_Delaunay_neighbor_lists (setup.py:45-99)    set_uniform_rays (:174-223)
  set_rays_spherical_symmetry (:226-339, uniform) set_Delaunay_boundary (:378-404)
  set_boundary_condition_zero / _CMB / _1D (:407-477)
  set_quadrature (:480-503)                       set_linedata_from_LAMDA_file (:632-779)

The LAMDA reader is this package's own (models.read_lamda); the reference reads
the files through astroquery, which is not a dependency here.  Species and
collision partners are looked up in model.chemistry.species.symbol with the
reference's name rules (p/o-H2 -> H2, e -> e-; an unknown partner gets the
dummy column 0).
"""
import numpy as np

from . import models


def _dimension(model):
    return model.parameters.dimension()


def _x_ordered_line(model):
    pos = np.array(model.geometry.points.position)
    if np.any(pos[:, 1:] != 0.0):
        raise ValueError("Not 1D (only x-coordinate can be non-zero).")
    x = pos[:, 0]
    if not (np.all(np.diff(x) >= 0) or np.all(np.diff(x) <= 0)):
        raise ValueError("Not 1D (x-coordinates are not ordered).")
    return x


def set_Delaunay_neighbor_lists(model):
    """Neighbours of every point: i-1 and i+1 along an ordered 1D line, the
    Delaunay (Voronoi) neighbours in 3D."""
    dim = _dimension(model)
    P = model.parameters.npoints()
    if dim == 1:
        _x_ordered_line(model)
        nbs = [[1]] + [[p - 1, p + 1] for p in range(1, P - 1)] + [[P - 2]]
        n_nbs = np.array([len(x) for x in nbs], dtype=np.uint32)
        flat = np.array([v for x in nbs for v in x], dtype=np.uint32)
    elif dim == 3:
        n_nbs, flat = models.delaunay_neighbors(np.array(model.geometry.points.position))
    elif dim == 2:
        raise ValueError("Dimension = 2 is not supported.")
    else:
        raise ValueError("Dimension should be 1 or 3.")
    model.geometry.points.neighbors.set(flat)
    model.geometry.points.n_neighbors.set(n_nbs)
    return model


def set_Delaunay_boundary(model):
    """The two end points of a 1D model (the reference supports 1D only here)."""
    if _dimension(model) != 1:
        raise ValueError("Dimension should be 1.")
    _x_ordered_line(model)
    model.parameters.set_nboundary(2)
    model.geometry.boundary.boundary2point.set(
        np.array([0, model.parameters.npoints() - 1], dtype=np.uint32))
    return model


def set_boundary_condition_zero(model):
    from .core import BoundaryCondition

    for b in range(model.parameters.nboundary()):
        model.geometry.boundary.set_boundary_condition(b, BoundaryCondition.Zero)
    return model


def set_boundary_condition_CMB(model):
    from .core import BoundaryCondition

    nb = model.parameters.nboundary()
    for b in range(nb):
        model.geometry.boundary.set_boundary_condition(b, BoundaryCondition.CMB)
    model.geometry.boundary.boundary_temperature.set(np.full(nb, models.T_CMB))
    return model


def set_boundary_condition_1D(model, T_in=models.T_CMB, T_out=models.T_CMB):
    from .core import BoundaryCondition

    if _dimension(model) != 1:
        raise ValueError("These boundary conditions only work for a 1D model.")
    if model.parameters.nboundary() != 2:
        raise ValueError("A 1D model must have exactly 2 boundary points.")
    for b in range(2):
        model.geometry.boundary.set_boundary_condition(b, BoundaryCondition.Thermal)
    model.geometry.boundary.boundary_temperature.set(np.array([T_in, T_out], dtype=np.float64))
    return model


def set_uniform_rays(model, randomize=False, first_ray=np.array([1.0, 0.0, 0.0])):
    """Uniform ray directions with equal weights: -x and +x in 1D (the spherically
    symmetric set when parameters.spherical_symmetry()), HEALPix directions rotated
    so that ray 0 points along first_ray in 3D."""
    nrays = model.parameters.nrays()
    dim = _dimension(model)
    if dim == 1:
        if model.parameters.spherical_symmetry():
            return set_rays_spherical_symmetry(model, uniform=True)
        if nrays != 2:
            raise ValueError("In 1D without spherical symmetry, nrays should be 2.")
        direction = np.array([[-1.0, 0.0, 0.0], [1.0, 0.0, 0.0]])
    elif dim == 3:
        if randomize:
            raise NotImplementedError("randomized ray sets: rotate the directions with "
                                      "models.random_rotations and set them directly")
        direction = models.uniform_rays(nrays, first_ray)
    elif dim == 2:
        raise ValueError("Dimension = 2 is not supported.")
    else:
        raise ValueError("Dimension should be 1 or 3.")
    model.geometry.rays.direction.set(direction)
    model.geometry.rays.weight.set(np.full(nrays, 1.0 / nrays))
    model.geometry.rays.use_adaptive_directions = False
    return model


def set_rays_spherical_symmetry(model, uniform=True, nextra=0, step=1):
    """The ray set of a spherically symmetric 1D model (uniform in angle)."""
    if not uniform or nextra != 0 or step != 1:
        raise NotImplementedError("only the uniform spherically symmetric ray set is built here")
    direction, weight = models.rays_spherical_symmetry(model.parameters.nrays())
    model.geometry.rays.direction.set(direction)
    model.geometry.rays.weight.set(weight)
    model.geometry.rays.use_adaptive_directions = False
    return model


def set_quadrature(model):
    """Gauss-Hermite roots and (normalised) weights of parameters.nquads() for every
    line-producing species."""
    roots, weights = models.quadrature(model.parameters.nquads())
    for l in range(model.parameters.nlspecs()):
        q = model.lines.lineProducingSpecies[l].quadrature
        q.roots.set(roots)
        q.weights.set(weights)
    return model


def _proper_name(name):
    if name in ("e",):
        return "e-"
    if name in ("pH2", "oH2", "p-H2", "o-H2", "PH2", "OH2"):
        return "H2"
    return name


def _species_number(symbols, name):
    proper = _proper_name(name)
    for i, s in enumerate(symbols):
        if s == proper:
            return i
    raise KeyError(f"Could not find species '{proper}' in the list of chemical species.")


def _molecule_name(path):
    with open(path) as fh:
        data = [l.strip() for l in fh if l.strip() and not l.strip().startswith("!")]
    return data[0].split()[0]


def set_linedata_from_LAMDA_file(model, fileNames, config={}):
    """The line data of every line-producing species from LAMDA files (one per
    species): levels, radiative transitions with the derived Einstein B and line
    frequencies, collision partners with the excitation rates from detailed
    balance; config['considered transitions'] keeps only the listed transitions."""
    from . import core

    if not isinstance(fileNames, list):
        fileNames = [fileNames]
    if len(fileNames) != model.parameters.nlspecs():
        raise ValueError("Number of provided LAMDA files != nlspecs")
    model.lines.resize_LineProducingSpecies(len(fileNames))
    symbols = list(model.chemistry.species.symbol)
    index = {}
    for code, (name, _) in models.LAMDA_PARTNERS.items():
        try:
            index[name] = _species_number(symbols, name)
        except KeyError:
            pass
    keep = config.get("considered transitions") if config else None
    if keep is not None and not isinstance(keep, list):
        keep = [keep]
    for lspec, path in enumerate(fileNames):
        sym = _molecule_name(path)
        num = _species_number(symbols, sym)
        d = models.read_lamda(path, num=num, species_index=index)
        ld = model.lines.lineProducingSpecies[lspec].linedata
        ld.sym = sym
        ld.num = num
        ld.inverse_mass = float(d["inverse_mass"])
        ld.nlev = int(d["nlev"])
        ld.energy = [float(x) for x in d["energy"]]
        ld.weight = [float(x) for x in d["weight"]]
        ks = list(range(int(d["nrad"]))) if not keep else [int(k) for k in keep]
        ld.nrad = len(ks)
        ld.irad = [int(d["irad"][k]) for k in ks]
        ld.jrad = [int(d["jrad"][k]) for k in ks]
        for f in ("A", "Bs", "Ba", "frequency"):
            setattr(ld, f, [float(d[f][k]) for k in ks])
        ld.colpar = core.VectorCollisionPartner()
        for c in d.get("colpar", []):
            cp = core.CollisionPartner()
            cp.num_col_partner = int(c["num_col_partner"])
            cp.orth_or_para_H2 = c["orth_or_para"]
            cp.icol, cp.jcol = [int(x) for x in c["icol"]], [int(x) for x in c["jcol"]]
            cp.tmp = [float(x) for x in c["tmp"]]
            cp.Cd = [float(x) for x in np.ravel(c["Cd"])]
            cp.Ce = [float(x) for x in np.ravel(c["Ce"])]
            ld.colpar.append(cp)
    return model
