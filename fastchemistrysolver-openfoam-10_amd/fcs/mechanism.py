"""Seeded synthetic mechanisms and cell fields for the chemistry benchmarks.

A mechanism is held in the flat description the C ABI takes
(``fcs_mech_desc`` in ``include/fcs.h``): NASA-7 thermo per species and a list
of reactions with integer, slot-expanded stoichiometry.  This mirrors what
``OptReaction::readInfo`` builds from the OpenFOAM dictionaries
(src/Reaction/OptReaction.C:348-1339): Arrhenius parameters, third-body
efficiency rows, Troe/Lindemann falloff parameters and PLOG tables.

Synthetic mechanisms are mass-balanced: every species is a composition over
four pseudo-elements carrying the atomic masses of H, C, O and N, and every
reaction conserves each element.  That keeps sum(Y) = 1 along trajectories,
so mass and enthalpy conservation can be used as invariants in the tests.
"""
from __future__ import annotations

import itertools
from dataclasses import dataclass, field

import numpy as np

# Universal gas constant [J/(kmol K)] and standard pressure [Pa]
# (src/Reaction/OptReaction.H:404,409).
RU = 8314.470066505449722171761095523834228515625
PSTD = 1.0e5

RXN_ARRHENIUS, RXN_THIRD_BODY, RXN_LINDEMANN, RXN_TROE, RXN_PLOG, RXN_SRI = 0, 1, 2, 3, 4, 5
FALLOFF = (RXN_LINDEMANN, RXN_TROE, RXN_SRI)
USES_M = (RXN_THIRD_BODY,) + FALLOFF
MAX_STOICH = 3     # reactant slots per reaction
MAX_SLOTS = 16     # product slots per reaction
ELEM_MASS = np.array([1.00794, 12.0107, 15.9994, 14.0067])  # H C O N [kg/kmol]

# Heats of formation are drawn in +-HF_RANGE J/mol.  At +-3e5 J/mol a random
# Dirichlet mixture of random species stores enough chemical energy to heat
# a cell by several thousand kelvin, past the NASA-7 fit range where the
# reference clamps T for rates but not for the energy equation; 27% of the
# 1,024 cells of the 10-species case then exceed maxSteps in the reference
# algorithm.  +-9e4 J/mol keeps every configuration inside the fit range.
HF_RANGE = 9.0e4
# Pre-exponential factors are drawn log-uniform in [1e6, 1e14] in the
# mol-cm^3-s units of CHEMKIN mechanism files with T^beta taken relative to
# 1000 K, then converted to the kmol-m^3-s, bare-T^beta form OpenFOAM reads:
# A_foam = A * (1e-3)^(order-1) * 1000^(-beta).
A_RANGE = (6.0, 14.0)
T_REF = 1000.0


@dataclass
class Reaction:
    rtype: int
    reversible: bool
    lhs: list
    rhs: list
    A: float
    beta: float
    Ta: float
    A0: float = 0.0
    beta0: float = 0.0
    Ta0: float = 0.0
    eff: np.ndarray | None = None
    troe: tuple = (0.0, 0.0, 0.0, 0.0)
    plog: tuple = ()      # (p[], A[], beta[], Ta[])
    sri: tuple = ()       # (a, b, c, d, e) of an SRI falloff
    rev_arr: tuple = ()   # (A, beta, Ta) of the reverse rate of a nonEquilibrium reaction


@dataclass
class Mechanism:
    names: list
    W: np.ndarray
    T_low: np.ndarray
    T_high: np.ndarray
    T_common: np.ndarray
    coeffs_low: np.ndarray    # [ns,7]
    coeffs_high: np.ndarray   # [ns,7]
    reactions: list = field(default_factory=list)
    composition: np.ndarray | None = None   # [ns,4] element counts

    @property
    def n_species(self) -> int:
        return len(self.names)

    @property
    def n_reactions(self) -> int:
        return len(self.reactions)

    def arrays(self) -> dict:
        """Flat, contiguous arrays matching ``fcs_mech_desc`` field by field."""
        ns, nr = self.n_species, self.n_reactions
        pmax = max([len(r.plog[0]) for r in self.reactions if r.rtype == RXN_PLOG] + [1])
        ss = max([MAX_STOICH] + [len(r.rhs) for r in self.reactions])
        a = dict(
            n_species=ns, n_reactions=nr, plog_max_points=pmax, slot_stride=ss,
            W=np.ascontiguousarray(self.W, np.float64),
            T_low=np.ascontiguousarray(self.T_low, np.float64),
            T_high=np.ascontiguousarray(self.T_high, np.float64),
            T_common=np.ascontiguousarray(self.T_common, np.float64),
            coeffs_high=np.ascontiguousarray(self.coeffs_high, np.float64).ravel(),
            coeffs_low=np.ascontiguousarray(self.coeffs_low, np.float64).ravel(),
            rxn_type=np.zeros(nr, np.int32), rxn_reversible=np.zeros(nr, np.int32),
            rxn_n_lhs=np.zeros(nr, np.int32), rxn_n_rhs=np.zeros(nr, np.int32),
            rxn_lhs=np.zeros(nr * ss, np.int32), rxn_rhs=np.zeros(nr * ss, np.int32),
            A=np.zeros(nr), beta=np.zeros(nr), Ta=np.zeros(nr),
            A0=np.zeros(nr), beta0=np.zeros(nr), Ta0=np.zeros(nr),
            efficiencies=np.ones(nr * ns), troe=np.zeros(nr * 4),
            plog_n=np.zeros(nr, np.int32), plog_p=np.zeros(nr * pmax),
            plog_A=np.zeros(nr * pmax), plog_beta=np.zeros(nr * pmax), plog_Ta=np.zeros(nr * pmax),
            sri=np.zeros(nr * 5), A_rev=np.zeros(nr), beta_rev=np.zeros(nr), Ta_rev=np.zeros(nr),
        )
        for r, rx in enumerate(self.reactions):
            a["rxn_type"][r] = rx.rtype
            a["rxn_reversible"][r] = 2 if rx.rev_arr else int(rx.reversible)
            if rx.rev_arr:
                a["A_rev"][r], a["beta_rev"][r], a["Ta_rev"][r] = rx.rev_arr
            if rx.rtype == RXN_SRI:
                a["sri"][r * 5:r * 5 + 5] = rx.sri
            a["rxn_n_lhs"][r] = len(rx.lhs)
            a["rxn_n_rhs"][r] = len(rx.rhs)
            a["rxn_lhs"][r * ss:r * ss + len(rx.lhs)] = rx.lhs
            a["rxn_rhs"][r * ss:r * ss + len(rx.rhs)] = rx.rhs
            a["A"][r], a["beta"][r], a["Ta"][r] = rx.A, rx.beta, rx.Ta
            a["A0"][r], a["beta0"][r], a["Ta0"][r] = rx.A0, rx.beta0, rx.Ta0
            if rx.eff is not None:
                a["efficiencies"][r * ns:(r + 1) * ns] = rx.eff
            a["troe"][r * 4:r * 4 + 4] = rx.troe
            if rx.rtype == RXN_PLOG:
                p, A, b, T = rx.plog
                k = len(p)
                a["plog_n"][r] = k
                a["plog_p"][r * pmax:r * pmax + k] = p
                a["plog_A"][r * pmax:r * pmax + k] = A
                a["plog_beta"][r * pmax:r * pmax + k] = b
                a["plog_Ta"][r * pmax:r * pmax + k] = T
        return a

    @staticmethod
    def from_arrays(a: dict, names=None) -> "Mechanism":
        """Inverse of arrays(): a Mechanism from the flat fcs_mech_desc arrays."""
        ns, nr = int(a["n_species"]), int(a["n_reactions"])
        pmax = int(a["plog_max_points"])
        ss = int(a["slot_stride"]) if "slot_stride" in a and int(a["slot_stride"]) > 0 else MAX_STOICH
        rx = []
        for r in range(nr):
            t = int(a["rxn_type"][r])
            x = Reaction(t, bool(a["rxn_reversible"][r]),
                         [int(v) for v in a["rxn_lhs"][r * ss:r * ss + int(a["rxn_n_lhs"][r])]],
                         [int(v) for v in a["rxn_rhs"][r * ss:r * ss + int(a["rxn_n_rhs"][r])]],
                         float(a["A"][r]), float(a["beta"][r]), float(a["Ta"][r]),
                         float(a["A0"][r]), float(a["beta0"][r]), float(a["Ta0"][r]))
            if t in USES_M:
                x.eff = np.array(a["efficiencies"][r * ns:(r + 1) * ns], np.float64)
            if "A_rev" in a and int(a["rxn_reversible"][r]) == 2:
                x.rev_arr = (float(a["A_rev"][r]), float(a["beta_rev"][r]), float(a["Ta_rev"][r]))
            if t == RXN_SRI:
                x.sri = tuple(float(v) for v in a["sri"][r * 5:r * 5 + 5])
            x.troe = tuple(float(v) for v in a["troe"][r * 4:r * 4 + 4])
            if t == RXN_PLOG:
                k = int(a["plog_n"][r])
                x.plog = tuple([float(v) for v in a[f"plog_{q}"][r * pmax:r * pmax + k]]
                               for q in ("p", "A", "beta", "Ta"))
            rx.append(x)
        return Mechanism(list(names) if names is not None else [f"S{i}" for i in range(ns)],
                         np.array(a["W"], np.float64), np.array(a["T_low"], np.float64),
                         np.array(a["T_high"], np.float64), np.array(a["T_common"], np.float64),
                         np.array(a["coeffs_low"], np.float64).reshape(ns, 7),
                         np.array(a["coeffs_high"], np.float64).reshape(ns, 7), rx)

    def stoich_matrix(self) -> np.ndarray:
        """Net stoichiometric coefficients nu[r, i] (products minus reactants)."""
        nu = np.zeros((self.n_reactions, self.n_species))
        for r, rx in enumerate(self.reactions):
            for s in rx.lhs:
                nu[r, s] -= 1
            for s in rx.rhs:
                nu[r, s] += 1
        return nu


# --------------------------------------------------------------------------
# thermo
# --------------------------------------------------------------------------
def _nasa7_for(rng, W):
    """NASA-7 low/high fits with cp/R rising smoothly from U[2.5,4] at 300 K
    to U[4,8] at 3000 K, hf ~ U(-HF_RANGE, HF_RANGE) J/mol, continuous at T_common."""
    clo = rng.uniform(2.5, 4.0)
    chi = rng.uniform(4.0, 8.0)
    # cp/R = clo + (chi-clo)*(2u-u^2), u = (T-300)/2700
    s = 1.0 / 2700.0
    u0 = -300.0 / 2700.0
    d = chi - clo
    # 2u - u^2 with u = s*T + u0
    c0 = clo + d * (2 * u0 - u0 * u0)
    c1 = d * (2 * s - 2 * u0 * s)
    c2 = d * (-s * s)
    low = np.array([c0, c1, c2, 0.0, 0.0, 0.0, 0.0])
    Tc = rng.uniform(800.0, 1200.0)
    # high fit: cp/R_low + e*((T-Tc)/1000)^2, same value and slope at Tc
    e = rng.uniform(-0.2, 0.2)
    high = low.copy()
    high[0] += e * Tc * Tc * 1e-6
    high[1] += -2.0 * e * Tc * 1e-6
    high[2] += e * 1e-6
    hf = rng.uniform(-HF_RANGE, HF_RANGE) * 1e3     # J/kmol
    s0 = rng.uniform(18.0, 32.0)                    # S/R at 298.15 K
    T0 = 298.15

    def h_over_r(a, T):
        return a[0] * T + a[1] * T**2 / 2 + a[2] * T**3 / 3 + a[3] * T**4 / 4 + a[4] * T**5 / 5

    def s_over_r(a, T):
        return a[0] * np.log(T) + a[1] * T + a[2] * T**2 / 2 + a[3] * T**3 / 3 + a[4] * T**4 / 4

    low[5] = hf / RU - h_over_r(low, T0)
    low[6] = s0 - s_over_r(low, T0)
    high[5] = h_over_r(low, Tc) + low[5] - h_over_r(high, Tc)
    high[6] = s_over_r(low, Tc) + low[6] - s_over_r(high, Tc)
    return low, high, Tc


# --------------------------------------------------------------------------
# species and balanced reactions
# --------------------------------------------------------------------------
def _pool():
    pool = []
    for c in itertools.product(range(0, 7), range(0, 4), range(0, 3), range(0, 3)):
        w = float(np.dot(c, ELEM_MASS))
        if 2.0 <= w <= 44.1:
            pool.append(c)
    return pool


def _species_set(rng, ns):
    pool = _pool()
    pool_set = set(pool)
    small = [c for c in pool if sum(c) <= 2]
    comps: list = []
    counts: dict = {}

    def add(c):
        if c in pool_set and counts.get(c, 0) < 2:
            comps.append(c)
            counts[c] = counts.get(c, 0) + 1
            return True
        return False

    while len(comps) < min(4, ns):
        add(small[rng.integers(len(small))])
    guard = 0
    while len(comps) < ns and guard < 100000:
        guard += 1
        r = rng.random()
        if r < 0.55:
            i, j = rng.integers(len(comps), size=2)
            c = tuple(int(x) for x in np.add(comps[i], comps[j]))
            if counts.get(c, 0) == 0:
                add(c)
        elif r < 0.85:
            c = comps[rng.integers(len(comps))]
            if sum(c) >= 2:
                sub = tuple(int(rng.integers(0, x + 1)) for x in c)
                rest = tuple(x - y for x, y in zip(c, sub))
                if sum(sub) > 0 and sum(rest) > 0:
                    cand = sub if rng.random() < 0.5 else rest
                    if counts.get(cand, 0) == 0:
                        add(cand)
        elif r < 0.93:
            c = comps[rng.integers(len(comps))]
            if counts.get(c, 0) == 1:
                add(c)      # isomer
        else:
            add(pool[rng.integers(len(pool))])
    return comps


def _balanced_reactions(comps):
    """All element-balanced reactions with 1-2 reactants and 1-2 products
    that share no species; each unordered pair reported once."""
    ns = len(comps)
    C = np.array(comps)
    groups: dict = {}
    sets = [(i,) for i in range(ns)] + [(i, j) for i in range(ns) for j in range(i, ns)]
    for s in sets:
        key = tuple(C[list(s)].sum(axis=0))
        groups.setdefault(key, []).append(s)
    out = []
    for g in groups.values():
        for a, b in itertools.combinations(g, 2):
            if set(a) & set(b):
                continue
            out.append((a, b))
    return out


def _arrhenius(rng, order):
    """(A, beta, Ta) in OpenFOAM units for a rate law of the given order."""
    A = 10 ** rng.uniform(*A_RANGE)
    beta = rng.uniform(-1.5, 1.5)
    Ta = rng.uniform(0.0, 25000.0)
    return A * 1e-3 ** (order - 1) * T_REF ** (-beta), beta, Ta


def synthetic_mechanism(n_species, n_reactions, n_third_body=0, n_troe=0, n_plog=0,
                        seed=1000, n_lindemann=0, n_sri=0, n_noneq=0) -> Mechanism:
    """Seeded synthetic mechanism as specified for the benchmark configurations:
    1-2 reactants and 1-2 products, reversible through Kc from the NASA-7 fits,
    A log-uniform in [1e6, 1e14], beta U[-1.5, 1.5], Ta U[0, 25000] K,
    third-body efficiencies U[0.5, 3], Troe alpha U[0.3, 0.9], T*** U[50, 300],
    T* U[500, 3000], T** U[3000, 8000] with k0 = 1e3 k_inf, PLOG with 5
    pressures log-spaced over 1e4..1e8 Pa.  Optionally n_sri SRI falloff
    reactions (a U[0.3, 1.5], b U[100, 1000], c U[500, 3000], d U[0.8, 1.2],
    e U[-0.3, 0.3]) and n_noneq Arrhenius / third-body reactions turned
    nonEquilibrium (reverse Arrhenius A 10^U[-1, 1] times the forward A, the
    forward beta, Ta U[0.5, 1.5] times the forward Ta); with both zero the
    mechanism is the same as without these options."""
    rng = np.random.default_rng(seed)
    for attempt in range(200):
        comps = _species_set(rng, n_species)
        cands = _balanced_reactions(comps)
        if len(comps) == n_species and len(cands) >= n_reactions:
            break
    else:
        raise RuntimeError("could not build a balanced species set")
    C = np.array(comps)
    W = C @ ELEM_MASS
    names = []
    seen: dict = {}
    for c in comps:
        base = "".join(f"{e}{k if k > 1 else ''}" for e, k in zip("HCON", c) if k > 0)
        seen[base] = seen.get(base, 0) + 1
        names.append(base if seen[base] == 1 else f"{base}_{seen[base]}")
    lows, highs, tcs = [], [], []
    for w in W:
        lo, hi, tc = _nasa7_for(rng, w)
        lows.append(lo)
        highs.append(hi)
        tcs.append(tc)
    mech = Mechanism(names=names, W=W, T_low=np.full(n_species, 200.0),
                     T_high=np.full(n_species, 3500.0), T_common=np.array(tcs),
                     coeffs_low=np.array(lows), coeffs_high=np.array(highs),
                     composition=C)
    # pick reactions: balance the four (n_lhs, n_rhs) size classes
    by_class: dict = {}
    for a, b in cands:
        if rng.random() < 0.5:
            a, b = b, a
        by_class.setdefault((len(a), len(b)), []).append((a, b))
    for v in by_class.values():
        rng.shuffle(v)
    classes = sorted(by_class)
    chosen = []
    while len(chosen) < n_reactions:
        k = classes[rng.integers(len(classes))]
        if by_class[k]:
            chosen.append(by_class[k].pop())
        elif not any(by_class.values()):
            break
    n_fall = n_troe + n_lindemann + n_sri
    types = ([RXN_THIRD_BODY] * n_third_body + [RXN_TROE] * n_troe
             + [RXN_LINDEMANN] * n_lindemann + [RXN_PLOG] * n_plog + [RXN_SRI] * n_sri)
    types += [RXN_ARRHENIUS] * (n_reactions - len(types))
    types = list(rng.permutation(types))
    assert n_fall + n_third_body + n_plog <= n_reactions
    for (lhs, rhs), t in zip(chosen, types):
        t = int(t)
        order = len(lhs) + (1 if t == RXN_THIRD_BODY else 0)
        A, b, Ta = _arrhenius(rng, order)
        rx = Reaction(rtype=t, reversible=True, lhs=list(lhs), rhs=list(rhs), A=A, beta=b, Ta=Ta)
        if t in USES_M:
            rx.eff = rng.uniform(0.5, 3.0, n_species)
        if t in FALLOFF:
            # k0 = 1e3 k_inf in CHEMKIN units; one order higher (the M factor)
            rx.A0, rx.beta0, rx.Ta0 = A * 1e3 * 1e-3, b, Ta
        if t == RXN_TROE:
            rx.troe = (rng.uniform(0.3, 0.9), rng.uniform(50, 300), rng.uniform(500, 3000),
                       rng.uniform(3000, 8000))
        if t == RXN_SRI:
            rx.sri = (rng.uniform(0.3, 1.5), rng.uniform(100, 1000), rng.uniform(500, 3000),
                      rng.uniform(0.8, 1.2), rng.uniform(-0.3, 0.3))
        if t == RXN_PLOG:
            p = np.logspace(4, 8, 5)
            pA = A * 10 ** rng.uniform(-1.0, 1.0, 5)
            pb = b + rng.uniform(-0.2, 0.2, 5)
            pT = np.clip(Ta + rng.uniform(-1000, 1000, 5), 0.0, None)
            rx.plog = (p, pA, pb, pT)
        mech.reactions.append(rx)
    if n_noneq:
        rng2 = np.random.default_rng(seed + 7)
        pool = [r for r in mech.reactions if r.rtype in (RXN_ARRHENIUS, RXN_THIRD_BODY)]
        for k in rng2.choice(len(pool), n_noneq, replace=False):
            r = pool[int(k)]
            r.rev_arr = (r.A * 10 ** rng2.uniform(-1.0, 1.0), r.beta, r.Ta * rng2.uniform(0.5, 1.5))
    return mech


# --------------------------------------------------------------------------
# fields
# --------------------------------------------------------------------------
@dataclass
class Field:
    T: np.ndarray        # [nc]
    p: np.ndarray        # [nc]
    Y: np.ndarray        # [ns, nc] species-major
    deltaTChem: np.ndarray
    cls: np.ndarray | None = None   # cell class of generated fields (counterflow: 0 cold, 1 product, 2 flame)

    @property
    def n_cells(self):
        return self.T.shape[0]

    def rho(self, mech: Mechanism) -> np.ndarray:
        """Ideal-gas density p / (Ru T sum(Y/W))."""
        return self.p / (RU * self.T * (self.Y / mech.W[:, None]).sum(axis=0))

    def subset(self, idx):
        return Field(self.T[idx].copy(), self.p[idx].copy(), self.Y[:, idx].copy(),
                     self.deltaTChem[idx].copy(), None if self.cls is None else self.cls[idx].copy())


def uniform_field(mech: Mechanism, n_cells, seed, T_range=(1000.0, 1500.0), p=101325.0,
                  dtchem0=1e-7) -> Field:
    """Cells with T ~ U(T_range), constant p and Y ~ Dirichlet(1)."""
    rng = np.random.default_rng(seed)
    T = rng.uniform(*T_range, n_cells)
    Y = rng.dirichlet(np.ones(mech.n_species), n_cells).T.copy()
    return Field(T, np.full(n_cells, p), Y, np.full(n_cells, dtchem0))


def mixture_fraction_field(mech: Mechanism, n_cells, seed, T_peak=2200.0, T_end=300.0,
                           Z_st=0.3, p=1.0e5, dtchem0=1e-7) -> Field:
    """1-D mixture-fraction profile: Z ~ U[0,1], Burke-Schumann-like T(Z)
    peaking at Z_st, Y linear in Z between two Dirichlet end compositions."""
    rng = np.random.default_rng(seed)
    Yox = rng.dirichlet(np.ones(mech.n_species))
    Yfu = rng.dirichlet(np.ones(mech.n_species))
    Z = rng.uniform(0.0, 1.0, n_cells)
    T = np.where(Z <= Z_st, T_end + (T_peak - T_end) * Z / Z_st,
                 T_end + (T_peak - T_end) * (1.0 - Z) / (1.0 - Z_st))
    Y = (1.0 - Z)[None, :] * Yox[:, None] + Z[None, :] * Yfu[:, None]
    return Field(T, np.full(n_cells, p), Y, np.full(n_cells, dtchem0))


def counterflow_field(mech: Mechanism, n_side, seed, equilibrate, T_react=500.0, n_base=16,
                      p=1.0e5, dtchem0=1e-7, n_rows=None) -> Field:
    """Synthetic 2-D counterflow field of n_side x n_rows cells (n_rows
    defaults to n_side), shuffled.

    Across the burner axis x in [0, 1] the cells fall into three bands: the
    two cold inflow streams |x - 0.5| > 0.275 (45 %, T in [300, 480) K, below
    T_react, oxidizer or fuel composition), the hot product layers
    0.075 < |x - 0.5| <= 0.275 (40 %, near-equilibrium states), and the thin
    flame front |x - 0.5| <= 0.075 (15 %, T from 900 K at the band edges to
    1800 K at the centre, composition linear in the mixture fraction).

    The product states come from ``equilibrate(field) -> field``, which must
    integrate a small Field of hot mixtures long enough to bring them close
    to chemical equilibrium (the GPU integrator does this at bench time);
    each product cell is one of n_base such states with 1e-9 relative noise
    on Y, so a 1 us flow step takes the integrator only a few sub-steps.
    """
    rng = np.random.default_rng(seed)
    rows = n_side if n_rows is None else n_rows
    n = n_side * rows
    Yox = rng.dirichlet(np.ones(mech.n_species))
    Yfu = rng.dirichlet(np.ones(mech.n_species))
    x = (np.arange(n) % n_side + 0.5) / n_side
    yy = (np.arange(n) // n_side + 0.5) / rows
    d = np.abs(x - 0.5)
    cold = d > 0.275
    flame = d <= 0.075
    hot = ~cold & ~flame
    Z = np.clip(x + 0.02 * np.sin(2 * np.pi * yy), 0.0, 1.0)
    T = np.empty(n)
    Y = np.empty((mech.n_species, n))
    # cold streams
    T[cold] = rng.uniform(300.0, 0.96 * T_react, cold.sum())
    Y[:, cold] = np.where(x[cold] < 0.5, Yox[:, None], Yfu[:, None])
    # flame front
    T[flame] = 1800.0 - 900.0 * d[flame] / 0.075 + rng.uniform(-5.0, 5.0, flame.sum())
    T[flame] = np.clip(T[flame], 900.0, 1800.0)
    Y[:, flame] = (1.0 - Z[flame])[None, :] * Yox[:, None] + Z[flame][None, :] * Yfu[:, None]
    # hot products: n_base equilibrated states
    Zb = np.linspace(0.2, 0.8, n_base)
    base = Field(np.linspace(1900.0, 2200.0, n_base), np.full(n_base, p),
                 (1.0 - Zb)[None, :] * Yox[:, None] + Zb[None, :] * Yfu[:, None], np.full(n_base, dtchem0))
    eq = equilibrate(base)
    k = np.minimum((yy[hot] * n_base).astype(int), n_base - 1)
    Yh = eq.Y[:, k] * (1.0 + 1e-9 * rng.standard_normal((mech.n_species, k.size)))
    Y[:, hot] = np.maximum(Yh, 0.0) / np.maximum(Yh, 0.0).sum(axis=0)
    T[hot] = eq.T[k]
    cls = np.where(cold, 0, np.where(flame, 2, 1)).astype(np.int32)
    order = rng.permutation(n)
    return Field(T[order].copy(), np.full(n, p), Y[:, order].copy(), np.full(n, dtchem0), cls[order].copy())


# Benchmark configurations (mechanism seed; field seed = mechanism seed + 1).
def config_mechanism(cfg: int) -> Mechanism:
    if cfg == 0:
        return synthetic_mechanism(10, 27, n_third_body=4, seed=1000)
    if cfg in (1, 2, 3):
        return synthetic_mechanism(53, 325, n_third_body=60, n_troe=25, n_plog=5, seed=1053)
    if cfg == 4:
        return synthetic_mechanism(160, 1200, n_third_body=200, n_troe=80, n_plog=20, seed=1160)
    raise ValueError(cfg)


def config_field(cfg: int, mech: Mechanism, n_cells=None, equilibrate=None) -> Field:
    """Cell field of a benchmark configuration.  Configurations 2 and 3 use
    the 512 x 512 counterflow field (n_cells must then be a square) and need
    an ``equilibrate`` callable (see counterflow_field)."""
    if cfg == 0:
        return uniform_field(mech, n_cells or 1024, seed=1001)
    if cfg == 1:
        return mixture_fraction_field(mech, n_cells or 65536, seed=1054)
    if cfg in (2, 3):
        n = n_cells or 262144
        side = int(round(np.sqrt(n)))
        if side * side != n:
            raise ValueError("configuration 2/3 fields are square: n_cells must be a square number")
        if equilibrate is None:
            raise ValueError("configuration 2/3 fields need an equilibrate callable")
        return counterflow_field(mech, side, seed=1054, equilibrate=equilibrate)
    if cfg == 4:
        # 512 x 256 counterflow field (131,072 cells) of the 160-species mechanism
        n = n_cells or 131072
        if n % 512:
            raise ValueError("configuration 4 fields have 512 cells per row: n_cells must be a multiple of 512")
        if equilibrate is None:
            raise ValueError("configuration 4 fields need an equilibrate callable")
        return counterflow_field(mech, 512, seed=1161, equilibrate=equilibrate, n_rows=n // 512)
    raise ValueError(cfg)


CONFIG_DT = {0: 1e-5, 1: 1e-6, 2: 1e-6, 3: 1e-6, 4: 1e-6}
