"""Seeded synthetic cells for the benchmark and parity configurations.

Every array comes from one ``numpy.random.Generator`` seeded by the caller.
Arrays are in the reference's cell-major (Fortran) layout: ``C[n, NSPEC]``
(variable then fixed species), ``RCONST[n, NREACT]``, ``ICNTRL[n, 20]``,
``RCNTRL[n, 20]``, ``Perm[n]`` (1-based), ``ATOL``/``RTOL[NVAR]``.

Rate constants.  GEOS-Chem's rate-law sources (gckpp_Rates.F90,
fullchem_RateLawFuncs.F90) and equation names are not part of the reference,
so thermal rate constants come from a documented synthetic rate law evaluated
at each cell's temperature T and air density M:

* a reaction's order is the number of V/F factors of its rate monomial;
* first-order reactions in the trailing block 856-1058 (1-based; the block
  that holds photolysis in GEOS-Chem's ordering) and the zero-order source
  A(884) = RCT(884)*F(3) are photolysis-like: J = J0 * max(cos SZA, 0) with J0
  log-uniform in [1e-8, 1e-2] s^-1 per reaction;
* other first-order reactions: k = k298 * exp(-E (1/T - 1/298)), k298
  log-uniform [1e-6, 1] s^-1, E uniform [0, 8000] K;
* second-order: k298 log-uniform [1e-16, 1e-10] cm3 s^-1, E uniform
  [-500, 3000] K; one in six of them is pressure dependent (termolecular
  with M folded in, as KPP does): k = k0 M / (1 + k0 M / kinf) with
  k0 = k0_300 (300/T)^m, k0_300 log-uniform [1e-32, 1e-29], m uniform [1, 4],
  kinf log-uniform [1e-12, 1e-10];
* third-order: k = k300 (300/T)^m, k300 log-uniform [1e-34, 1e-30].

Reactions whose rate is a literal in the mechanism (A(126), A(735) of
fullchem) do not read RCONST; their entry is set to 0.

Fixed species are taken as (H2O, N2, O2) = (r M, 0.7808 M, 0.2095 M), the
order inferred from the mechanism's O1D + F(1..3) reactions (not pinned by
any reference file).
"""
from __future__ import annotations

import numpy as np

from .mechanism import Mechanism

KB = 1.380649e-23  # J/K
PHOTOLYSIS_FIRST = 855  # 0-based index of reaction 856


def _logu(rng, lo, hi, size=None):
    return np.exp(rng.uniform(np.log(lo), np.log(hi), size))


def reaction_classes(mech: Mechanism):
    """Per reaction: 'lit' (literal rate), 'photo', 'first', 'second', 'third'."""
    out = []
    for r, f in enumerate(mech.ir["rate"]):
        if f[0][0] == "lit":
            out.append("lit")
            continue
        order = sum(1 for k, _ in f[1:] if k in ("V", "F"))
        if (r >= PHOTOLYSIS_FIRST and order <= 1) or order == 0:
            out.append("photo")
        elif order == 1:
            out.append("first")
        elif order == 2:
            out.append("second")
        else:
            out.append("third")
    return out


def rate_constants(mech: Mechanism, rng, T, M, cos_sza):
    """RCONST[n, NREACT] for cells with temperature T [K], air density M
    [molec cm-3] and cos(SZA)."""
    cls = reaction_classes(mech)
    n, R = T.shape[0], mech.nreact
    out = np.zeros((n, R))
    inv = (1.0 / T - 1.0 / 298.0)[:, None]
    t300 = (300.0 / T)[:, None]
    idx = {c: np.array([r for r, k in enumerate(cls) if k == c], dtype=np.int64)
           for c in ("photo", "first", "second", "third")}
    # photolysis-like
    r = idx["photo"]
    J0 = _logu(rng, 1e-8, 1e-2, r.size)
    out[:, r] = J0[None, :] * np.maximum(cos_sza, 0.0)[:, None]
    # thermal first order
    r = idx["first"]
    k298, E = _logu(rng, 1e-6, 1.0, r.size), rng.uniform(0.0, 8000.0, r.size)
    out[:, r] = k298[None, :] * np.exp(-E[None, :] * inv)
    # second order, one in six pressure dependent
    r = idx["second"]
    k298, E = _logu(rng, 1e-16, 1e-10, r.size), rng.uniform(-500.0, 3000.0, r.size)
    kb = k298[None, :] * np.exp(-E[None, :] * inv)
    pdep = rng.uniform(size=r.size) < 1.0 / 6.0
    k0 = _logu(rng, 1e-32, 1e-29, r.size)[None, :] * t300 ** rng.uniform(1.0, 4.0, r.size)[None, :]
    kinf = _logu(rng, 1e-12, 1e-10, r.size)[None, :]
    k0M = k0 * M[:, None]
    kp = k0M / (1.0 + k0M / kinf)
    out[:, r] = np.where(pdep[None, :], kp, kb)
    # third order
    r = idx["third"]
    k300, m = _logu(rng, 1e-34, 1e-30, r.size), rng.uniform(0.0, 4.0, r.size)
    out[:, r] = k300[None, :] * t300 ** m[None, :]
    return out


def _controls(n, nvar, atol=1e-2, rtol=0.5e-2):
    ICNTRL = np.zeros((n, 20), np.int32)
    ICNTRL[:, 0] = 1    # autonomous
    ICNTRL[:, 1] = 0    # vector tolerances
    ICNTRL[:, 2] = 4    # Rodas3
    ICNTRL[:, 14] = -1  # no Update_* calls
    RCNTRL = np.zeros((n, 20))
    return (ICNTRL, RCNTRL, np.full(nvar, atol), np.full(nvar, rtol))


def _assemble(mech, rng, T, P_hpa, cos_sza, n):
    M = (P_hpa * 100.0) / (KB * T) * 1e-6  # molec cm-3
    h2o = _logu(rng, 1e-6, 3e-2, n) * M
    C = np.empty((n, mech.nspec))
    C[:, :mech.nvar] = _logu(rng, 1e-2, 1e12, (n, mech.nvar))
    fix = np.stack([h2o, 0.7808 * M, 0.2095 * M], axis=1)
    C[:, mech.nvar:] = fix[:, :mech.nfix] if mech.nfix <= 3 else np.pad(
        fix, ((0, 0), (0, mech.nfix - 3)), constant_values=1e10)
    RCONST = rate_constants(mech, rng, T, M, cos_sza)
    ICNTRL, RCNTRL, ATOL, RTOL = _controls(n, mech.nvar)
    return dict(C=C, RCONST=RCONST, ICNTRL=ICNTRL, RCNTRL=RCNTRL, ATOL=ATOL, RTOL=RTOL,
                Perm=np.arange(1, n + 1, dtype=np.int32), Tstart=0.0, Tend=600.0,
                T=T, P_hpa=P_hpa, cos_sza=cos_sza)


def random_cells(mech: Mechanism, n: int = 2048, seed: int = 0) -> dict:
    """Independent cells: T ~ U(200,310) K, P ~ logU(10,1000) hPa, SZA ~ U(0,120)
    deg, variable species logU(1e-2,1e12), dt = 600 s."""
    rng = np.random.default_rng(seed)
    T = rng.uniform(200.0, 310.0, n)
    P = _logu(rng, 10.0, 1000.0, n)
    sza = np.deg2rad(rng.uniform(0.0, 120.0, n))
    return _assemble(mech, rng, T, P, np.cos(sza), n)


def std_atmosphere_T(z_km):
    """US standard atmosphere 1976 temperature [K] at geopotential height z [km]."""
    z = np.asarray(z_km, dtype=np.float64)
    layers = [(0, 288.15, -6.5), (11, 216.65, 0.0), (20, 216.65, 1.0), (32, 228.65, 2.8),
              (47, 270.65, 0.0), (51, 270.65, -2.8), (71, 214.65, -2.0), (86, 184.65, 0.0)]
    T = np.empty_like(z)
    for k, (z0, t0, lapse) in enumerate(layers):
        z1 = layers[k + 1][0] if k + 1 < len(layers) else np.inf
        m = (z >= z0) & (z < z1)
        T[m] = t0 + lapse * (z[m] - z0)
    T[z < 0] = 288.15 - 6.5 * z[z < 0]
    return T


def grid_cells(mech: Mechanism, nlon: int, nlat: int, nlev: int, seed: int = 1) -> dict:
    """Cells of a lon x lat x level grid (lon fastest, as GEOS-Chem's I,J,L):
    pressure log-spaced 1000 -> 0.01 hPa over the levels, standard-atmosphere
    temperature by level with +-15 K uniform noise, SZA from the cell's
    longitude/latitude at a random UTC hour and day of year."""
    rng = np.random.default_rng(seed)
    n = nlon * nlat * nlev
    lon = -180.0 + (360.0 / nlon) * np.arange(nlon)
    lat = np.linspace(-89.0, 89.0, nlat)
    plev = np.exp(np.linspace(np.log(1000.0), np.log(0.01), nlev))
    L, J, I = np.meshgrid(np.arange(nlev), np.arange(nlat), np.arange(nlon), indexing="ij")
    I, J, L = I.ravel(), J.ravel(), L.ravel()
    P = plev[L]
    z = -7.0 * np.log(P / 1013.25)
    T = std_atmosphere_T(z) + rng.uniform(-15.0, 15.0, n)
    utc = rng.uniform(0.0, 24.0)
    doy = rng.uniform(1.0, 365.0)
    decl = np.deg2rad(23.44) * np.sin(2.0 * np.pi * (284.0 + doy) / 365.0)
    hour_angle = np.deg2rad((utc + lon[I] / 15.0 - 12.0) * 15.0)
    la = np.deg2rad(lat[J])
    cos_sza = np.sin(la) * np.sin(decl) + np.cos(la) * np.cos(decl) * np.cos(hour_angle)
    d = _assemble(mech, rng, T, P, cos_sza, n)
    d["grid"] = (nlon, nlat, nlev)
    return d


def grid_4x5_cells(mech: Mechanism, seed: int = 1) -> dict:
    """4x5 grid: 72 x 46 columns x 72 levels = 238,464 cells."""
    return grid_cells(mech, 72, 46, 72, seed)


def grid_2x25_cells(mech: Mechanism, seed: int = 2) -> dict:
    """2x2.5 grid: 144 x 91 columns x 72 levels = 943,488 cells."""
    return grid_cells(mech, 144, 91, 72, seed)


def adjoint_cells(mech: Mechanism, n: int = 65536, seed: int = 3) -> dict:
    """Adjoint workload: n cells taken evenly across the 4x5 grid generator's
    238,464 cells (all levels and latitudes represented), plus the loss
    weights ``w[NVAR]`` ~ N(0, 1) per species, all from `seed`."""
    d = grid_cells(mech, 72, 46, 72, seed)
    idx = np.linspace(0, d["C"].shape[0] - 1, n).round().astype(np.int64)
    out = subset(d, idx)
    out["w"] = np.random.default_rng(seed + 1000).standard_normal(mech.nvar)
    return out


def subset(d: dict, idx) -> dict:
    """Cells `idx` of a generated batch (identity launch order)."""
    idx = np.asarray(idx)
    out = dict(d)
    for k in ("C", "RCONST", "ICNTRL", "RCNTRL", "T", "P_hpa", "cos_sza"):
        if k in d:
            out[k] = np.ascontiguousarray(d[k][idx])
    out["Perm"] = np.arange(1, idx.size + 1, dtype=np.int32)
    out.pop("grid", None)
    return out


def cost_order(nsteps, descending: bool = True):
    """1-based launch order sorted by a per-cell cost key (previous Nsteps):
    a stable counting sort, as the model's pre-scan does
    (src/geos-chem/GeosCore/fullchem_mod.F90:905-981, descending there;
    standalone/kpp_batch.F90:847-880 sorts ascending)."""
    key = np.asarray(nsteps, dtype=np.int64)
    order = np.argsort(-key if descending else key, kind="stable")
    return (order + 1).astype(np.int32)


def stress_cells(mech: Mechanism, n: int, seed: int = 4, max_steps: int = 1000) -> dict:
    """Cells of a synthetic stress mechanism: concentrations log-uniform
    [1e-2, 1e10]; rate constants log-uniform [1e-6, 1e4] s^-1 in scaled units,
    i.e. RCONST = k / (1e6)^(order-1) so that k is the pseudo-first-order rate
    at a concentration of 1e6; dt = 600 s; fixed species 1e10.  The random
    initial states give a stiff transient (median ~400 accepted steps, a few
    cells in a thousand need 10^4), so Max_no_steps = ICNTRL(4) is set to
    `max_steps`: cells beyond it end with IERR = -6 for the caller's CPU retry,
    as GEOS-Chem handles failed cells."""
    rng = np.random.default_rng(seed)
    C = np.empty((n, mech.nspec))
    C[:, :mech.nvar] = _logu(rng, 1e-2, 1e10, (n, mech.nvar))
    C[:, mech.nvar:] = 1e10
    order = np.array([sum(1 for k, _ in f[1:] if k in ("V", "F")) for f in mech.ir["rate"]])
    k = _logu(rng, 1e-6, 1e4, (n, mech.nreact))
    RCONST = k / (1e6 ** np.maximum(order - 1, 0))[None, :]
    ICNTRL, RCNTRL, ATOL, RTOL = _controls(n, mech.nvar)
    ICNTRL[:, 3] = max_steps
    return dict(C=C, RCONST=RCONST, ICNTRL=ICNTRL, RCNTRL=RCNTRL, ATOL=ATOL, RTOL=RTOL,
                Perm=np.arange(1, n + 1, dtype=np.int32), Tstart=0.0, Tend=600.0)
