"""OpenFOAM-10 dictionary reader and mechanism loader.

Reads the `reactions` and `speciesThermo` dictionaries the plug-in takes
(the reaction forms OptReaction::readInfo accepts, src/Reaction/OptReaction.C:
348-1175: (ir)reversibleArrhenius, (ir)reversibleThirdBodyArrhenius,
(ir)reversibleArrheniusLindemannFallOff, (ir)reversibleArrheniusTroeFallOff,
(ir)reversibleArrheniusSRIFallOff, (ir)reversibleArrheniusPLOG (also read as
(ir)reversiblePlog), nonEquilibriumReversible(ThirdBody)Arrhenius) into a
fcs.mechanism.Mechanism, and volScalarField files
(`internalField uniform v` or `nonuniform List<scalar> n (...)`) into arrays.

The dictionary grammar handled is the subset these files use: `key value;`
entries, `name { ... }` sub-dictionaries, `( ... )` lists optionally preceded
by their length, uniform lists `N{value}`, double-quoted strings, and words that may contain balanced
parentheses (species names such as CH2(S)).  Comments are // and /* */.
"""
from __future__ import annotations

import re

import numpy as np

from .mechanism import (RXN_ARRHENIUS, RXN_LINDEMANN, RXN_PLOG, RXN_SRI, RXN_THIRD_BODY, RXN_TROE,
                        MAX_SLOTS, MAX_STOICH, Mechanism, Reaction)

_PUNCT = set("{};")


def _strip_comments(text: str) -> str:
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
    return re.sub(r"//[^\n]*", " ", text)


def tokenize(text: str):
    """Tokens: '{', '}', ';', '(', ')', strings (as ('str', s)) and words."""
    text = _strip_comments(text)
    out, i, n = [], 0, len(text)
    while i < n:
        c = text[i]
        if c.isspace():
            i += 1
        elif c in _PUNCT:
            out.append(c)
            i += 1
        elif c == '"':
            j = text.index('"', i + 1)
            out.append(("str", text[i + 1:j]))
            i = j + 1
        elif c in "()":
            out.append(c)
            i += 1
        else:
            # a word: stops at whitespace, ; { } " and at an unbalanced ')';
            # '(' inside a word opens a nested group (CH2(S))
            j, depth = i, 0
            while j < n:
                d = text[j]
                if d.isspace() or d in _PUNCT or d == '"':
                    break
                if d == "(":
                    if j == i:
                        break
                    depth += 1
                elif d == ")":
                    if depth == 0:
                        break
                    depth -= 1
                j += 1
            out.append(text[i:j])
            i = j
    return out


def _number(w):
    try:
        return float(w)
    except (TypeError, ValueError):
        return None


class _Parser:
    def __init__(self, tokens):
        self.t = tokens
        self.i = 0

    def peek(self):
        return self.t[self.i] if self.i < len(self.t) else None

    def take(self):
        tok = self.t[self.i]
        self.i += 1
        return tok

    def value(self):
        """One value: a list, a dict, a string, a number or a word."""
        tok = self.take()
        if tok == "(":
            return self.list_body()
        if tok == "{":
            return self.dict_body()
        if isinstance(tok, tuple):
            return tok[1]
        num = _number(tok)
        if num is not None and self.peek() == "(" and float(num).is_integer():
            self.take()
            return self.list_body()      # counted list: "53 ( ... )"
        if num is not None and self.peek() == "{" and float(num).is_integer():
            self.take()                  # uniform list: "7{0}" is seven zeros
            v = self.value()
            if self.take() != "}":
                raise ValueError("malformed uniform list")
            return [v] * int(num)
        return num if num is not None else tok

    def list_body(self):
        items = []
        while self.peek() != ")":
            items.append(self.value())
        self.take()
        return items

    def dict_body(self, top=False):
        d = {}
        while True:
            tok = self.peek()
            if tok is None:
                if top:
                    return d
                raise ValueError("unterminated dictionary")
            if tok == "}":
                self.take()
                return d
            if tok == ";":
                self.take()
                continue
            key = self.take()
            key = key[1] if isinstance(key, tuple) else key
            if self.peek() == "{":
                self.take()
                d[key] = self.dict_body()
                continue
            vals = []
            while self.peek() != ";":
                vals.append(self.value())
            self.take()
            d[key] = vals[0] if len(vals) == 1 else vals


def parse(text: str) -> dict:
    """An OpenFOAM dictionary file as nested Python dicts / lists."""
    p = _Parser(tokenize(text))
    d = p.dict_body(top=True)
    d.pop("FoamFile", None)
    return d


def read(path: str) -> dict:
    with open(path) as f:
        return parse(f.read())


# --------------------------------------------------------------------------
# mechanism
# --------------------------------------------------------------------------
_FORMS = {
    "Arrhenius": RXN_ARRHENIUS,
    "ThirdBodyArrhenius": RXN_THIRD_BODY,
    "ArrheniusLindemannFallOff": RXN_LINDEMANN,
    "ArrheniusTroeFallOff": RXN_TROE,
    "ArrheniusSRIFallOff": RXN_SRI,
    "Plog": RXN_PLOG,
    "ArrheniusPLOG": RXN_PLOG,   # OptReaction.C:1129-1166 (rows in ArrheniusData)
}
_NONEQ = "nonEquilibriumReversible"


def _side(text, index):
    """Species slots of one side of a reaction equation: integer
    coefficients expand to repeated slots (OptReaction.C:99-126)."""
    slots = []
    for term in text.split("+"):
        term = term.strip()
        if not term:
            continue
        m = re.match(r"^(\d+)\s*(.+)$", term)
        coef, name = (int(m.group(1)), m.group(2).strip()) if m else (1, term)
        if "^" in name:
            raise ValueError(f"reaction orders are not supported: {term!r}")
        if name not in index:
            raise ValueError(f"unknown species {name!r}")
        slots += [index[name]] * coef
    if not 1 <= len(slots) <= MAX_SLOTS:
        raise ValueError(f"{len(slots)} species slots on one side of a reaction (1..{MAX_SLOTS} supported)")
    return slots


def _efficiencies(entry, index, ns, default=1.0):
    eff = np.full(ns, default)
    for name, val in entry:
        if name not in index:
            raise ValueError(f"unknown species {name!r} in efficiencies")
        eff[index[name]] = float(val)
    return eff


def _arrhenius(d):
    return float(d["A"]), float(d["beta"]), float(d["Ta"])


def mechanism(reactions: dict, thermo: dict) -> Mechanism:
    """Mechanism from parsed `reactions` and `speciesThermo` dictionaries
    (species order of the `species` list of the reactions file, else of the
    thermo file; a thermo database may hold species the mechanism does not
    use)."""
    names = list(reactions.get("species") or thermo.get("species"))
    index = {n: i for i, n in enumerate(names)}
    ns = len(names)
    W = np.zeros(ns)
    Tl, Th, Tc = np.zeros(ns), np.zeros(ns), np.zeros(ns)
    lo, hi = np.zeros((ns, 7)), np.zeros((ns, 7))
    for i, n in enumerate(names):
        sp = thermo[n]
        W[i] = float(sp["specie"]["molWeight"])
        th = sp["thermodynamics"]
        Tl[i], Th[i], Tc[i] = float(th["Tlow"]), float(th["Thigh"]), float(th["Tcommon"])
        hi[i] = [float(v) for v in th["highCpCoeffs"]]
        lo[i] = [float(v) for v in th["lowCpCoeffs"]]
    rxns = []
    for key, r in reactions["reactions"].items():
        typ = r["type"]
        noneq = typ.startswith(_NONEQ)
        rev = not typ.startswith("irreversible")
        if noneq:
            form = typ[len(_NONEQ):]
            if form not in ("Arrhenius", "ThirdBodyArrhenius"):
                raise ValueError(f"{key}: reaction type {typ!r} is not supported")
        else:
            form = typ[len("irreversible"):] if typ.startswith("irreversible") else typ[len("reversible"):]
        if form not in _FORMS:
            raise ValueError(f"{key}: reaction type {typ!r} is not supported")
        t = _FORMS[form]
        lhs_s, rhs_s = r["reaction"].split("=", 1)
        lhs, rhs = _side(lhs_s, index), _side(rhs_s, index)
        # reactant sides of up to 3 slots (product sides up to MAX_SLOTS, _side)
        if len(lhs) > MAX_STOICH:
            raise ValueError(f"{key}: {typ} {r['reaction']!r}: {len(lhs)} reactant slots; reactant sides "
                             f"take at most {MAX_STOICH}")
        rx = Reaction(t, rev, lhs, rhs, 0.0, 0.0, 0.0)
        if noneq:
            # forward / reverse sub-dictionaries; third-body efficiencies of
            # both directions come from the forward one (OptReaction.C:617-700)
            rx.A, rx.beta, rx.Ta = _arrhenius(r["forward"])
            rx.rev_arr = _arrhenius(r["reverse"])
            if t == RXN_THIRD_BODY:
                rx.eff = _efficiencies(r["forward"]["coeffs"], index, ns)
        elif t in (RXN_ARRHENIUS, RXN_THIRD_BODY):
            rx.A, rx.beta, rx.Ta = _arrhenius(r)
        if t == RXN_THIRD_BODY and not noneq:
            rx.eff = _efficiencies(r["coeffs"], index, ns)
        if t in (RXN_LINDEMANN, RXN_TROE, RXN_SRI):
            rx.A, rx.beta, rx.Ta = _arrhenius(r["kInf"])
            rx.A0, rx.beta0, rx.Ta0 = _arrhenius(r["k0"])
            rx.eff = _efficiencies(r["thirdBodyEfficiencies"]["coeffs"], index, ns)
        if t == RXN_TROE:
            F = r["F"]
            rx.troe = (float(F["alpha"]), float(F["Tsss"]), float(F["Ts"]), float(F["Tss"]))
        if t == RXN_SRI:
            F = r["F"]
            rx.sri = tuple(float(F[k]) for k in "abcde")
        if t == RXN_PLOG:
            rows = r["ArrheniusData"] if "ArrheniusData" in r else r["coeffs"]
            p = [float(row[0]) for row in rows]
            A = [float(row[1]) for row in rows]
            b = [float(row[2]) for row in rows]
            Ta = [float(row[3]) for row in rows]
            rx.plog = (p, A, b, Ta)
        rxns.append(rx)
    return Mechanism(names, W, Tl, Th, Tc, lo, hi, rxns)


def load_mechanism(reactions_path: str, thermo_path: str) -> Mechanism:
    return mechanism(read(reactions_path), read(thermo_path))


def read_field(path: str, n_cells: int | None = None) -> np.ndarray:
    """internalField of a volScalarField file."""
    d = read(path)
    f = d["internalField"]
    if isinstance(f, list) and f and f[0] == "uniform":
        if n_cells is None:
            raise ValueError(f"{path}: uniform field needs n_cells")
        return np.full(n_cells, float(f[1]))
    if isinstance(f, list) and f[0] == "nonuniform":
        vals = f[-1]
        return np.asarray(vals, np.float64)
    raise ValueError(f"{path}: unsupported internalField")
