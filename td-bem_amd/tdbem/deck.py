"""Reader and writer for the reference's input-deck format (src/InputT.cpp:84-994).

Format rules kept from the reference: delimiters " \\t=;," (:21); one keyword per line; node,
element and dof ids are 1-based in the file and 0-based in memory (:529-534, :747); dof id =
3*node + direction; numbers are read with atoi/atof semantics (a leading numeric prefix, so
"5." reads as 5 for an integer field); an unknown keyword is an error (:135-136).
"""
import re
from dataclasses import dataclass, field

import numpy as np

_DELIM = re.compile(r"[ \t=;,]+")
_INT = re.compile(r"\s*[+-]?\d+")
_FLOAT = re.compile(r"\s*[+-]?(\d+\.?\d*([eE][+-]?\d+)?|\.\d+([eE][+-]?\d+)?)")


def _atoi(s):
    m = _INT.match(s)
    return int(m.group(0)) if m else 0


def _atof(s):
    m = _FLOAT.match(s)
    return float(m.group(0)) if m else 0.0


def _tokens(line):
    return [t for t in _DELIM.split(line.strip()) if t]


class DeckError(ValueError):
    """Raised where the reference throws a const char* (its message is kept)."""


@dataclass
class Deck:
    formulation: int = 1
    numstep: int = 1
    dt: float = 0.0
    interpolation: int = 1
    nq: int = 6
    E: float = 1.0
    nu: float = 0.0
    density: float = 1.0
    sd: int = 3
    if_ext: int = 0
    nnd: int = 0
    etype: int = 1
    nel: int = 0
    coords: np.ndarray = None
    ien: np.ndarray = None
    ubc_dof: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))
    ubc_val: np.ndarray = field(default_factory=lambda: np.zeros(0))
    fbc_dof: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))
    fbc_val: np.ndarray = field(default_factory=lambda: np.zeros(0))
    interface: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))
    # second region of formulation 3 (src/InputT.cpp:548-693); read, not used by the hot paths
    nnd_aux: int = 0
    etype_aux: int = 1
    ennd_aux: int = 4
    nel_aux: int = 0
    coords_aux: np.ndarray = None
    ien_aux: np.ndarray = None

    @property
    def ennd(self):
        return 4 if self.etype == 1 else 8


_SCALARS = {
    "formulation": ("formulation", _atoi), "numstep": ("numstep", _atoi),
    "dt": ("dt", _atof), "interpolation": ("interpolation", lambda s: int(_atof(s))),
    "nq": ("nq", lambda s: int(_atof(s))), "E": ("E", _atof), "nu": ("nu", _atof),
    "density": ("density", _atof), "sd": ("sd", _atoi), "if_ext": ("if_ext", _atoi),
    "nnd": ("nnd", _atoi), "nel": ("nel", _atoi),
}


def read_deck(path):
    """Parse a deck the way InputT::ReadInput does (formulation 3 aux blocks are not read)."""
    d = Deck()
    with open(path) as f:
        lines = f.read().split("\n")
    counts = {"UBC_num": 0, "FBC_num": 0, "num_interface": 0}
    i = 0

    def take_list(n, conv):
        nonlocal i
        out = []
        while len(out) < n:
            if i >= len(lines):
                raise DeckError("unexpected end of file in a list")
            toks = _tokens(lines[i])
            i += 1
            if not toks:
                toks = [""]
            for t in toks:
                if len(out) < n:
                    out.append(conv(t))
        return out

    while i < len(lines):
        toks = _tokens(lines[i])
        i += 1
        if not toks:
            continue
        key = toks[0]
        if key in _SCALARS:
            if len(toks) < 2:
                raise DeckError(f"{key} is not set properly")
            name, conv = _SCALARS[key]
            setattr(d, name, conv(toks[1]))
        elif key == "EType":
            if len(toks) < 2:
                raise DeckError("number of nodes is not set properly")
            d.etype = _atoi(toks[1])
            if d.etype not in (1, 2):
                raise DeckError("InputT::ReadInput, unsupported element type")
        elif key == "nodes":
            rows = []
            for _ in range(d.nnd):
                t = _tokens(lines[i]); i += 1
                rows.append([_atof(t[k]) if k < len(t) else 0.0 for k in range(3)])
            d.coords = np.array(rows, dtype=np.float64).reshape(-1, 3)
        elif key in ("IEN", "IEN_aux") and (key == "IEN" or d.formulation == 3):
            aux = key == "IEN_aux"
            nel, ennd = (d.nel_aux, d.ennd_aux) if aux else (d.nel, d.ennd)
            rows = []
            for _ in range(nel):
                t = _tokens(lines[i]) if i < len(lines) else []
                i += 1
                if len(t) < ennd:
                    # the reference passes strtok's NULL to atoi here (src/InputT.cpp:531-535)
                    raise DeckError(f"{key} row {len(rows) + 1} has fewer than {ennd} node ids")
                rows.append([_atoi(t[k]) - 1 for k in range(ennd)])
            arr = np.array(rows, dtype=np.int32).reshape(-1, ennd)
            if aux:
                d.ien_aux = arr
            else:
                d.ien = arr
        elif d.formulation == 3 and key in ("nnd_aux", "nel_aux"):
            if len(toks) < 2:
                raise DeckError("number of nodes is not set properly")
            setattr(d, key, _atoi(toks[1]))
        elif d.formulation == 3 and key == "etype_aux":
            if len(toks) < 2:
                raise DeckError("number of nodes is not set properly")
            d.etype_aux = _atoi(toks[1])
            # src/InputT.cpp:595-597 tests fEType (not fEType_Aux) for the 8-node case
            if d.etype_aux == 1:
                d.ennd_aux = 4
            elif d.etype == 2:
                d.ennd_aux = 8
            else:
                raise DeckError("InputT::ReadInput, unsupported element type")
        elif d.formulation == 3 and key == "nodes_aux":
            rows = []
            for _ in range(d.nnd_aux):
                t = _tokens(lines[i]); i += 1
                rows.append([_atof(t[k]) if k < len(t) else 0.0 for k in range(3)])
            d.coords_aux = np.array(rows, dtype=np.float64).reshape(-1, 3)
        elif key in counts:
            if len(toks) < 2:
                raise DeckError(f"{key} is not set properly")
            counts[key] = _atoi(toks[1])
        elif key == "UBC_dof":
            d.ubc_dof = np.array([v - 1 for v in take_list(counts["UBC_num"], _atoi)], np.int32)
        elif key == "UBC_values":
            d.ubc_val = np.array(take_list(counts["UBC_num"], _atof), np.float64)
        elif key == "FBC_dof":
            d.fbc_dof = np.array([v - 1 for v in take_list(counts["FBC_num"], _atoi)], np.int32)
        elif key == "FBC_values":
            d.fbc_val = np.array(take_list(counts["FBC_num"], _atof), np.float64)
        elif key == "interface":
            d.interface = np.array([v - 1 for v in take_list(counts["num_interface"], _atoi)],
                                   np.int32)
        else:
            raise DeckError("InputT::ReadInput, unaccepted key word")
    return d


def _num(v):
    return format(float(v), ".17g")


def write_deck(path, mesh_coords, ien, *, E, nu, density, dt, numstep, nq=6, if_ext=0,
               ubc_dof=(), ubc_val=(), fbc_dof=(), fbc_val=(), etype=1, interpolation=4,
               formulation=1, interface=None):
    """Write a deck in the reference format (full-precision numbers, 1-based ids); interface:
    0-based interface node ids of a two-region deck (num_interface / interface blocks)."""
    ubc_dof = np.asarray(ubc_dof, dtype=np.int64)
    fbc_dof = np.asarray(fbc_dof, dtype=np.int64)
    out = [f"formulation={formulation};", f"numstep={numstep};", f"dt={_num(dt)};",
           f"interpolation={interpolation};", f"nq={nq};", "",
           f"E={_num(E)};", f"nu={_num(nu)};", f"density={_num(density)};", "",
           "sd=3;", f"nnd={len(mesh_coords)};", f"if_ext={if_ext};", f"EType={etype};",
           f"nel={len(ien)};", "", "nodes="]
    out += [f"{_num(x)} {_num(y)} {_num(z)}" for x, y, z in np.asarray(mesh_coords, np.float64)]
    out += ["", "IEN="]
    out += [",".join(str(int(v) + 1) for v in row) for row in np.asarray(ien)]

    def lst(vals, fmt):
        vals = list(vals)
        rows = [",".join(fmt(v) for v in vals[k:k + 10]) for k in range(0, len(vals), 10)]
        return rows if rows else [""]

    out += ["", f"UBC_num={len(ubc_dof)};", "UBC_dof="] + lst(ubc_dof + 1, str)
    out += ["", "UBC_values="] + lst(ubc_val, _num)
    out += ["", f"FBC_num={len(fbc_dof)};", "FBC_dof="] + lst(fbc_dof + 1, str)
    out += ["", "FBC_values="] + lst(fbc_val, _num)
    if interface is not None:
        iface = np.asarray(interface, dtype=np.int64)
        out += ["", f"num_interface={len(iface)}", "", "interface="] + lst(iface + 1, str)
    with open(path, "w") as f:
        f.write("\n".join(out) + "\n")
