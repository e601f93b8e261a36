"""Reader of Kadath's binary space-and-fields files (the `.dat` half of a FUKA
solution), big-endian as written by fwrite_be (src/Utilities/fwrite_be.cpp:25).

The layout is the chain of `save` routines, read here in the same order:

  Space_adapted_bh::save (space_adapted_bh.cpp:99-105): nbr_domains, ndim, type_base,
      then each domain's save
  Space_spheric::save: the same header, then the domains
  Domain_nucleus::save (domain_nucleus.cpp:52-59): nbr_points, nbr_coefs (Dim_array),
      ndim, type_base, center (Point), alpha
  Domain_shell::save (domain_shell.cpp:53-61): ..., center, alpha, beta
  Domain_compact::save (domain_compact.cpp:58-65): ..., center, alpha
  Domain_shell_outer_adapted::save (domain_shell_outer_adapted.cpp:444-452):
      ..., center, inner_radius (double), outer_radius (Val_domain)
  Domain_shell_inner_adapted::save (domain_shell_inner_adapted.cpp:447-455):
      ..., center, outer_radius (double), inner_radius (Val_domain)
  Dim_array::save (dim_array.cpp:73-76): ndim, sizes
  Point::save (point.cpp:51-54): ndim, coordinates
  Array<T>::save (array.cpp:83-87): Dim_array, count, data; the element of index
      (i0, i1, i2) sits at (i0 d1 + i1) d2 + i2 (array.cpp:103-111): the LAST index
      fastest, i.e. phi, then theta, then r for a domain array
  Base_spectral::save (base_spectral.cpp:70-77): 0 if defined, ndim, the 1-D basis
      arrays (Array<int>) when defined
  Val_domain::save (val_domain.cpp:83-95): base, is_zero, 0 + conf array when held in
      configuration space (else 1), 0 + coef array when held in coefficients (else 1)
  Scalar::save (scalar.cpp:64-67): one Val_domain per domain
  Tensor::save (tensor.cpp:361-369): Base_tensor (Array<int>), index types
      (Array<int>), valence, n_comp, each component's Scalar

A FUKA solution file is the space followed by its fields
(bco_utils::save_to_file, bco_utilities.hpp:78-88): for a black hole (bh_3d_xcts)
conf, lapse (Scalars) and shift (Vector).

write_solution writes the same chain back (Space::save, Scalar::save, Tensor::save), so
a file read and written again is identical byte for byte.

Host-side file handling only: the arrays read here are handed to a system's
set_coef / set_field.
"""
import numpy as np

# domain kinds of a Space_adapted_bh (space_adapted_bh.cpp:59-73): nucleus, shell with an
# adapted (homothetic) outer boundary, shell with an adapted inner boundary, shells, compact
NUCLEUS, SHELL_OUTER_ADAPTED, SHELL_INNER_ADAPTED, SHELL, COMPACT = range(5)


class _Stream:
    def __init__(self, data):
        self.b = memoryview(data)
        self.p = 0

    def ints(self, n):
        v = np.frombuffer(self.b, dtype=">i4", count=n, offset=self.p).astype(np.int64)
        self.p += 4 * n
        return v

    def int(self):
        return int(self.ints(1)[0])

    def doubles(self, n):
        v = np.frombuffer(self.b, dtype=">f8", count=n, offset=self.p).astype(np.float64)
        self.p += 8 * n
        return v

    def double(self):
        return float(self.doubles(1)[0])

    def dim_array(self):
        return tuple(int(x) for x in self.ints(self.int()))

    def array(self, kind):
        dims = self.dim_array()
        n = self.int()
        if n != int(np.prod(dims)):
            raise ValueError(f"array of {dims} holds {n} values")
        data = self.ints(n) if kind == "i" else self.doubles(n)
        return data, dims

    def point(self):
        return self.doubles(self.int())


def r_fastest(data, dims):
    """A domain array as stored (phi fastest, dims (n_r, n_theta, n_phi)) in this
    package's order: r fastest, then theta, then phi (flat)."""
    return np.ascontiguousarray(np.asarray(data).reshape(dims).transpose(2, 1, 0)).ravel()


class ValDomain:
    """One Val_domain: spectral bases (one int array per dimension), and the values in
    configuration space and/or coefficient space as stored (phi index fastest), or zero.
    conf_values() / coef_values() give them r index fastest."""

    def __init__(self, bases, is_zero, conf, coef, conf_dims, coef_dims, ndim=3):
        self.bases, self.is_zero, self.ndim = bases, is_zero, ndim
        self.conf, self.coef = conf, coef
        self.conf_dims, self.coef_dims = conf_dims, coef_dims

    def conf_values(self):
        return None if self.conf is None else r_fastest(self.conf, self.conf_dims)

    def coef_values(self):
        return None if self.coef is None else r_fastest(self.coef, self.coef_dims)

    def theta_base(self, k):
        """Basis id of the theta lines of phi coefficient k (bases[1] is indexed by k)."""
        return int(self.bases[1][0][k]) if self.bases is not None else None


def _val_domain(st):
    defined = st.int() == 0
    ndim = st.int()
    bases = [st.array("i") for _ in range(ndim)] if defined else None
    is_zero = st.int() == 1
    conf = coef = None
    conf_dims = coef_dims = None
    if st.int() == 0:
        conf, conf_dims = st.array("d")
    if st.int() == 0:
        coef, coef_dims = st.array("d")
    return ValDomain(bases, is_zero, conf, coef, conf_dims, coef_dims, ndim)


class Domain:
    def __init__(self, kind, nbr_points, nbr_coefs, type_base, center, **mapping):
        self.kind, self.nbr_points, self.nbr_coefs = kind, nbr_points, nbr_coefs
        self.type_base, self.center = type_base, center
        self.__dict__.update(mapping)

    def bounds(self):
        """(r_in, r_out) of the domain (r_out = inf for the compactified domain)."""
        if self.kind == NUCLEUS:
            return 0.0, self.alpha
        if self.kind == SHELL:
            return self.beta - self.alpha, self.beta + self.alpha
        if self.kind == COMPACT:
            return -0.5 / self.alpha, np.inf
        if self.kind == SHELL_OUTER_ADAPTED:
            return self.inner_radius, _constant(self.outer_radius)
        return _constant(self.inner_radius), self.outer_radius


def _constant(vd):
    """The value of a radius held as a Val_domain that is constant over the sphere (a
    homothetic domain's); raises if it varies."""
    v = vd.conf if vd.conf is not None else None
    if v is None:
        c = vd.coef
        if c is None or np.abs(c[1:]).max() > 1e-13 * abs(c[0]):
            raise ValueError("adapted radius is not a constant")
        return float(c[0])
    if np.ptp(v) > 1e-13 * abs(v.max()):
        raise ValueError("adapted radius is not a constant")
    return float(v.mean())


def _domain(st, kind):
    nbr_points = st.dim_array()
    nbr_coefs = st.dim_array()
    ndim = st.int()
    if ndim != 3:
        raise ValueError(f"{ndim}-dimensional domain")
    type_base = st.int()
    center = st.point()
    if kind == NUCLEUS:
        return Domain(kind, nbr_points, nbr_coefs, type_base, center, alpha=st.double())
    if kind == SHELL:
        alpha = st.double()
        return Domain(kind, nbr_points, nbr_coefs, type_base, center, alpha=alpha, beta=st.double())
    if kind == COMPACT:
        return Domain(kind, nbr_points, nbr_coefs, type_base, center, alpha=st.double())
    if kind == SHELL_OUTER_ADAPTED:
        rin = st.double()
        return Domain(kind, nbr_points, nbr_coefs, type_base, center, inner_radius=rin, outer_radius=_val_domain(st))
    rout = st.double()
    return Domain(kind, nbr_points, nbr_coefs, type_base, center, outer_radius=rout, inner_radius=_val_domain(st))


class Space:
    def __init__(self, kind, ndim, type_base, domains):
        self.kind, self.ndim, self.type_base, self.domains = kind, ndim, type_base, domains

    def radii(self):
        """Domain boundaries r_0 < r_1 < ... (the nucleus radius first)."""
        return [d.bounds()[1] for d in self.domains[:-1]]


def _space(st, kind):
    ndom, ndim, type_base = st.int(), st.int(), st.int()
    if kind == "adapted_bh":
        # Space_adapted_bh(FILE*) (space_adapted_bh.cpp:59-75)
        kinds = [NUCLEUS, SHELL_OUTER_ADAPTED, SHELL_INNER_ADAPTED] + [SHELL] * (ndom - 4) + [COMPACT]
    elif kind == "spheric":
        kinds = [NUCLEUS] + [SHELL] * (ndom - 2) + [COMPACT]
    else:
        raise ValueError(f"unknown space kind {kind}")
    return Space(kind, ndim, type_base, [_domain(st, k) for k in kinds])


def _scalar(st, ndom):
    return [_val_domain(st) for _ in range(ndom)]


def _tensor(st, ndom):
    basis = st.array("i")
    types = st.array("i")
    valence = st.int()
    n_comp = st.int()
    return {"basis": basis, "index_types": types, "valence": valence,
            "components": [_scalar(st, ndom) for _ in range(n_comp)]}


def read_solution(path, space_kind="adapted_bh", fields=("scalar", "scalar", "vector")):
    """Read a space and the fields saved after it.  fields: the kinds in file order
    ("scalar" or "vector"); a FUKA black hole holds conf, lapse, shift.  Returns
    (Space, [per field: list of ValDomain (scalar) or tensor dict (vector)])."""
    with open(path, "rb") as f:
        data = f.read()
    st = _Stream(data)
    sp = _space(st, space_kind)
    ndom = len(sp.domains)
    out = []
    for k in fields:
        out.append(_scalar(st, ndom) if k == "scalar" else _tensor(st, ndom))
    if st.p != len(data):
        raise ValueError(f"{len(data) - st.p} trailing bytes after the fields")
    return sp, out


class _Writer:
    def __init__(self):
        self.parts = []

    def ints(self, v):
        self.parts.append(np.asarray(v, dtype=">i4").tobytes())

    def int(self, x):
        self.ints([int(x)])

    def doubles(self, v):
        self.parts.append(np.asarray(v, dtype=">f8").tobytes())

    def double(self, x):
        self.doubles([float(x)])

    def dim_array(self, dims):
        self.int(len(dims))
        self.ints(dims)

    def array(self, data, dims, kind):
        self.dim_array(dims)
        self.int(len(data))
        (self.ints if kind == "i" else self.doubles)(data)


def _w_val_domain(w, v):
    w.int(0 if v.bases is not None else 1)
    w.int(v.ndim)
    for data, dims in v.bases or []:
        w.array(data, dims, "i")
    w.int(1 if v.is_zero else 0)
    for data, dims in ((v.conf, v.conf_dims), (v.coef, v.coef_dims)):
        if data is None:
            w.int(1)
        else:
            w.int(0)
            w.array(data, dims, "d")


def _w_domain(w, d):
    w.dim_array(d.nbr_points)
    w.dim_array(d.nbr_coefs)
    w.int(3)
    w.int(d.type_base)
    w.int(len(d.center))
    w.doubles(d.center)
    if d.kind in (NUCLEUS, COMPACT):
        w.double(d.alpha)
    elif d.kind == SHELL:
        w.double(d.alpha)
        w.double(d.beta)
    elif d.kind == SHELL_OUTER_ADAPTED:
        w.double(d.inner_radius)
        _w_val_domain(w, d.outer_radius)
    else:
        w.double(d.outer_radius)
        _w_val_domain(w, d.inner_radius)


def write_solution(path, space, fields):
    """Write a space and its fields (as returned by read_solution) in the same format."""
    w = _Writer()
    w.int(len(space.domains))
    w.int(space.ndim)
    w.int(space.type_base)
    for d in space.domains:
        _w_domain(w, d)
    for f in fields:
        if isinstance(f, dict):
            w.array(*f["basis"], "i")
            w.array(*f["index_types"], "i")
            w.int(f["valence"])
            w.int(len(f["components"]))
            for comp in f["components"]:
                for v in comp:
                    _w_val_domain(w, v)
        else:
            for v in f:
                _w_val_domain(w, v)
    with open(path, "wb") as fh:
        fh.write(b"".join(w.parts))


def read_info(path):
    """The `.info` steering file (Kadath's Configurator text format: `name value` lines
    in `{ }` sections, configurator_boost_imp.cpp:48-127), as {section: {name: value}};
    numeric values as floats, `on`/`off` as bools."""
    out, stack = {}, []
    with open(path) as f:
        for line in f:
            s = line.strip()
            if not s:
                continue
            if s == "{":
                continue
            if s == "}":
                if stack:
                    stack.pop()
                continue
            parts = s.split()
            if len(parts) == 1:
                stack.append(parts[0])
                out.setdefault(parts[0], {})
                continue
            name, val = parts[0], " ".join(parts[1:])
            if val in ("on", "off"):
                v = val == "on"
            else:
                try:
                    v = float(val)
                except ValueError:
                    v = val
            out.setdefault(stack[-1] if stack else "", {})[name] = v
    return out
