"""Run parameters: the `donnees.f` namelist of the reference and its mapping
onto the C ABI's cf_params.

Mirrors global_data.cuf:31-94 (variables, defaults and the namelist group
DONNEES_NAMELIST), lec_donnees (maillage.cuf:168-194) and the string-option
codes of data_transfer.cuf:43-66.
"""
from __future__ import annotations

import dataclasses
import re
import struct
from typing import Any

from ._lib import CF_MAX_HYDRO, CF_MAX_IO, CfParams


def f32(x: float) -> float:
    """A Fortran default-real literal widened to double."""
    return struct.unpack("f", struct.pack("f", x))[0]


INLET_CODES = {"transm": 1, "inflow": 2, "phys_inflow": 3}
OUTLET_CODES = {"fixedheight": 1, "transm": 2}


@dataclasses.dataclass
class Params:
    """Namelist variables of DONNEES_NAMELIST with the reference defaults
    (global_data.cuf:31-62).  Variables without a default there (iflux,
    tolisec, cfl, ts, inlet) get a value that must be overridden."""

    gp: float = 9.81
    is_override_manning: int = 0
    override_manning: float = 0.0
    is_cgns: int = 0
    meshfile_path: str = "./"
    meshfile: str = ""
    elt_bound: int = 0
    nombre_entrees: int = 1
    nombre_sorties: int = 1
    eqlin_barrage: int = 0
    x1eqbar: float = 0.0
    x2eqbar: float = 0.0
    y1eqbar: float = 0.0
    y2eqbar: float = 0.0
    h_amont: float = 0.0
    u_amont: float = 0.0
    v_amont: float = 0.0
    h_aval: float = 0.0
    u_aval: float = 0.0
    v_aval: float = 0.0
    plan: int = 0
    xplan: float = 0.0
    yplan: float = 0.0
    zplan: float = 1.0
    xpoint: float = 0.0
    ypoint: float = 0.0
    zpoint: float = 0.0
    solinit: int = 0
    fich_sol_init: str = ""
    inlet: str = "inflow"
    inlet_name: list = dataclasses.field(default_factory=lambda: [""] * CF_MAX_IO)
    debitglob: list = dataclasses.field(default_factory=lambda: [0.0] * CF_MAX_IO)
    debit_var: int = 0
    loi_debitglob: str = ""
    outlet: str = "fixedheight"
    outlet_name: list = dataclasses.field(default_factory=lambda: [""] * CF_MAX_IO)
    h_sortie: list = dataclasses.field(default_factory=lambda: [0.0] * CF_MAX_IO)
    iflux: int = 1
    ilimiteur: int = 1
    iupwind: int = 1
    tolisec: float = 1.0e-5
    friction: int = 0
    fricimplic: int = 1
    diffusion: int = 0
    ts: float = 1.0
    cfl: float = 0.9
    tol_reg_perm: float = 0.0
    is_dt_constant: int = 0
    constant_dt: float = 0.0
    freqaffich: int = 1000
    tolaffiche: float = dataclasses.field(default_factory=lambda: f32(1e-2))
    sol_z_offset: float = 0.0
    solrestart: int = 0
    solvisu: int = 0
    visu_snapshots: int = 0
    sortie_finale_bluekenue: int = 0
    # not in the namelist group, initialised at declaration (global_data.cuf:59)
    tol: float = dataclasses.field(default_factory=lambda: f32(1e-6))
    # inflow hydrographs (this build's extension): {inlet number: (t[], Q[])}
    hydrographs: dict = dataclasses.field(default_factory=dict)

    def inlet_code(self) -> int:
        key = self.inlet.strip().lower()
        if key not in INLET_CODES:
            raise ValueError("Bad inlet !")  # data_transfer.cuf:52-54
        return INLET_CODES[key]

    def outlet_code(self) -> int:
        key = self.outlet.strip().lower()
        if key not in OUTLET_CODES:
            raise ValueError("Bad outlet !")  # data_transfer.cuf:63-65
        return OUTLET_CODES[key]

    def to_cf(self, long_entree=None, ts: float | None = None) -> CfParams:
        """cf_params for the C ABI.  long_entree: inlet lengths (from
        cf_inlet_lengths); ts overrides the end time (1e300 = run a fixed
        number of steps)."""
        p = CfParams()
        p.gp = self.gp
        p.tolisec = self.tolisec
        p.tol = self.tol
        p.cfl = self.cfl
        p.ts = self.ts if ts is None else ts
        p.constant_dt = self.constant_dt
        p.override_manning = self.override_manning
        p.tol_reg_perm = self.tol_reg_perm
        p.iflux = self.iflux
        p.ilimiteur = self.ilimiteur
        p.friction = self.friction
        p.fricimplic = self.fricimplic
        p.inlet = self.inlet_code()
        p.outlet = self.outlet_code()
        p.is_dt_constant = self.is_dt_constant
        p.is_override_manning = self.is_override_manning
        p.nombre_entrees = self.nombre_entrees
        for k in range(CF_MAX_IO):
            p.debitglob[k] = float(self.debitglob[k])
            p.h_sortie[k] = float(self.h_sortie[k])
            p.long_entree[k] = float(long_entree[k]) if long_entree is not None and k < len(long_entree) else 0.0
        for inlet, (t, q) in self.hydrographs.items():
            k = int(inlet) - 1
            if not (0 <= k < CF_MAX_IO) or len(t) != len(q) or len(t) > CF_MAX_HYDRO:
                raise ValueError(f"bad hydrograph for inlet {inlet}")
            p.hydro_n[k] = len(t)
            for i, (ti, qi) in enumerate(zip(t, q)):
                p.hydro_t[k * CF_MAX_HYDRO + i] = float(ti)
                p.hydro_q[k * CF_MAX_HYDRO + i] = float(qi)
        return p


# ---------------------------------------------------------------------------
# Fortran namelist reader for `donnees.f`

_ARRAY_FIELDS = {"debitglob", "h_sortie", "inlet_name", "outlet_name"}


def _strip_comment(line: str) -> str:
    out, quote = [], None
    for ch in line:
        if quote:
            out.append(ch)
            if ch == quote:
                quote = None
        elif ch in "'\"":
            quote = ch
            out.append(ch)
        elif ch == "!":
            break
        else:
            out.append(ch)
    return "".join(out)


def _tokens(s: str) -> list[str]:
    """Split a namelist value list on commas/blanks, keeping quoted strings."""
    toks, cur, quote, i = [], [], None, 0
    while i < len(s):
        ch = s[i]
        if quote:
            if ch == quote:
                if i + 1 < len(s) and s[i + 1] == quote:  # doubled quote
                    cur.append(ch)
                    i += 2
                    continue
                quote = None
                toks.append("".join(cur))
                cur = []
            else:
                cur.append(ch)
        elif ch in "'\"":
            quote = ch
            cur = ["\x00"]  # marks a string token
        elif ch in ", \t\n":
            if cur:
                toks.append("".join(cur))
                cur = []
        else:
            cur.append(ch)
        i += 1
    if cur:
        toks.append("".join(cur))
    return toks


def _convert(tok: str) -> Any:
    if tok.startswith("\x00"):
        return tok[1:]
    low = tok.lower()
    if low in (".true.", "t", ".t."):
        return True
    if low in (".false.", "f", ".f."):
        return False
    try:
        return int(tok)
    except ValueError:
        pass
    return float(low.replace("d", "e"))


def parse_namelist(text: str, group: str = "donnees_namelist") -> dict[str, Any]:
    """Parse one namelist group; returns {name: value or [values]}.
    Array elements may be given as name=v1,v2 or name(i)=v; r*v repeats."""
    body = "\n".join(_strip_comment(l) for l in text.splitlines())
    m = re.search(r"&\s*" + group + r"\b", body, re.I)
    if not m:
        raise ValueError(f"namelist group &{group} not found")
    # the group ends at the first '/' outside a quoted string
    quote, end = None, None
    for pos in range(m.end(), len(body)):
        ch = body[pos]
        if quote:
            if ch == quote:
                quote = None
        elif ch in "'\"":
            quote = ch
        elif ch == "/":
            end = pos
            break
    if end is None:
        raise ValueError(f"namelist group &{group} is not terminated by '/'")
    body = body[m.end():end]
    keys = list(re.finditer(r"([A-Za-z_]\w*)\s*(\(\s*(\d+)\s*\))?\s*=", body))
    out: dict[str, Any] = {}
    for n, km in enumerate(keys):
        end = keys[n + 1].start() if n + 1 < len(keys) else len(body)
        vals: list[Any] = []
        for tok in _tokens(body[km.end():end]):
            if re.fullmatch(r"\d+\*.+", tok):
                r, v = tok.split("*", 1)
                vals += [_convert(v)] * int(r)
            else:
                vals.append(_convert(tok))
        name = km.group(1).lower()
        if km.group(3):
            idx = int(km.group(3)) - 1
            cur = out.setdefault(name, {})
            if isinstance(cur, dict):
                for k, v in enumerate(vals):
                    cur[idx + k] = v
        else:
            out[name] = vals if len(vals) != 1 else vals[0]
    return out


def read_donnees(path: str = "donnees.f") -> Params:
    """Read `donnees.f` like lec_donnees (maillage.cuf:168-194)."""
    with open(path, "r", encoding="latin-1") as fh:
        nml = parse_namelist(fh.read())
    return params_from_namelist(nml)


def params_from_namelist(nml: dict[str, Any]) -> Params:
    p = Params()
    fields = {f.name for f in dataclasses.fields(Params)}
    for name, val in nml.items():
        if name not in fields:
            raise ValueError(f"unknown namelist variable {name!r}")
        if name in _ARRAY_FIELDS:
            arr = list(getattr(p, name))
            if isinstance(val, dict):
                for k, v in val.items():
                    arr[k] = v
            else:
                vals = val if isinstance(val, list) else [val]
                arr[: len(vals)] = vals
            setattr(p, name, arr)
        else:
            if isinstance(val, list):
                val = val[0]
            cur = getattr(p, name)
            if isinstance(cur, bool) or isinstance(val, bool):
                val = int(bool(val))
            elif isinstance(cur, float):
                val = float(val)
            elif isinstance(cur, int):
                val = int(val)
            elif isinstance(cur, str):
                val = str(val)
            setattr(p, name, val)
    p.inlet = p.inlet.strip()
    return p
