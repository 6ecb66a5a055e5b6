"""The reference's input files as cases: input.nml (read_input,
src/param.f90:86-178) and spheres.in (read_sphere_file,
src/two_fluid.f90:438-470).

`parse(text)` parses the &dns, &scalar and &two_fluid groups of a
Fortran namelist with the reference's declarations and defaults (bounds,
types, array sections such as `cbcvel(0:1,1:3,1) = 'D','D', ...`, repeat
counts `3*0.`, logicals T/F/.true./.false., d exponents, ! comments);
`case_from_input(text, spheres_text, vof)` turns the result into a case
dict of `fizzy_amd.cases` (the derived values of src/param.f90:170-178
follow from it in the library). Groups the library does not use (&scalar,
&cudecomp) are read or skipped as the reference does; settings outside the
hot path that would change a run (an initial velocity other than 'zer', a
wall-turbulence flag) are refused.
"""
import re

# name: (group, type, bounds) -- src/param.f90:13-50; type i integer, r real,
# l logical, s string; bounds [] for a scalar, else [(lo, hi), ...]
B3 = [(1, 3)]
BC2 = [(0, 1), (1, 3)]
BC3 = [(0, 1), (1, 3), (1, 3)]
DECL = {
    "ng": ("dns", "i", B3), "l": ("dns", "r", B3), "gtype": ("dns", "i", []), "gr": ("dns", "r", []),
    "cfl": ("dns", "r", []), "dtmax": ("dns", "r", []), "dt_f": ("dns", "r", []),
    "is_solve_ns": ("dns", "l", []), "is_track_interface": ("dns", "l", []), "inivel": ("dns", "s", []),
    "is_wallturb": ("dns", "l", []), "is_forced_hit": ("dns", "l", []), "nstep": ("dns", "i", []),
    "time_max": ("dns", "r", []), "tw_max": ("dns", "r", []), "stop_type": ("dns", "l", B3),
    "restart": ("dns", "l", []), "is_overwrite_save": ("dns", "l", []), "nsaves_max": ("dns", "i", []),
    "icheck": ("dns", "i", []), "iout0d": ("dns", "i", []), "iout1d": ("dns", "i", []),
    "iout2d": ("dns", "i", []), "iout3d": ("dns", "i", []), "isave": ("dns", "i", []),
    "cbcvel": ("dns", "s", BC3), "cbcpre": ("dns", "s", BC2), "bcvel": ("dns", "r", BC3),
    "bcpre": ("dns", "r", BC2), "bforce": ("dns", "r", B3), "gacc": ("dns", "r", B3),
    "dims": ("dns", "i", [(1, 2)]),
    "inisca": ("scalar", "s", []), "cbcsca": ("scalar", "s", BC2), "bcsca": ("scalar", "r", BC2),
    "ssource": ("scalar", "r", []),
    "inipsi": ("two_fluid", "s", []), "cbcpsi": ("two_fluid", "s", BC2), "cbcnor": ("two_fluid", "s", BC3),
    "bcnor": ("two_fluid", "r", BC3), "bcpsi": ("two_fluid", "r", BC2), "sigma": ("two_fluid", "r", []),
    "rho12": ("two_fluid", "r", [(1, 2)]), "mu12": ("two_fluid", "r", [(1, 2)]),
    "ka12": ("two_fluid", "r", [(1, 2)]), "cp12": ("two_fluid", "r", [(1, 2)]),
    "beta12": ("two_fluid", "r", [(1, 2)]), "psi_thickness_factor": ("two_fluid", "r", []),
}
GROUPS = ("dns", "scalar", "two_fluid")


def _size(bounds):
    n = 1
    for lo, hi in bounds:
        n *= hi - lo + 1
    return n


def defaults(vof=False):
    """The values read_input sets before it reads the file (src/param.f90:123-151)."""
    d = dict(ng=[128, 128, 128], l=[1.0, 1.0, 1.0], gtype=1, gr=0.0, cfl=0.95, dtmax=1.0e9, dt_f=-1.0,
             is_solve_ns=True, is_track_interface=True, inivel="zer", is_wallturb=False, is_forced_hit=False,
             nstep=1000, time_max=1.0, tw_max=0.5, stop_type=[True, False, False], restart=False,
             is_overwrite_save=True, nsaves_max=0, icheck=10, iout0d=10, iout1d=100, iout2d=1000, iout3d=500,
             isave=1000, cbcvel=["P"] * 18, cbcpre=["P"] * 6, bcvel=[0.0] * 18, bcpre=[0.0] * 6,
             bforce=[0.0] * 3, gacc=[0.0] * 3, dims=[0, 0], inisca="zer", cbcsca=["P"] * 6, bcsca=[0.0] * 6,
             ssource=0.0, inipsi="uni", cbcpsi=["P"] * 6, cbcnor=["P"] * 18, bcpsi=[0.0] * 6, bcnor=[0.0] * 18,
             sigma=0.0, rho12=[1.0, 1.0], mu12=[0.01, 0.01], ka12=[0.01, 0.01], cp12=[1.0, 1.0],
             beta12=[1.0, 1.0], psi_thickness_factor=0.50 if vof else 0.51)
    return d


def _strip_comments(text):
    out = []
    for line in text.splitlines():
        q, keep = None, []
        for ch in line:
            if q:
                keep.append(ch)
                if ch == q:
                    q = None
            elif ch in "'\"":
                q = ch
                keep.append(ch)
            elif ch == "!":
                break
            else:
                keep.append(ch)
        out.append("".join(keep))
    return "\n".join(out)


def _tokens(body):
    """Quoted strings, and runs of everything else split on commas and blanks."""
    toks, i, n = [], 0, len(body)
    while i < n:
        ch = body[i]
        if ch in "'\"":
            j = i + 1
            s = []
            while j < n:
                if body[j] == ch:
                    if j + 1 < n and body[j + 1] == ch:  # doubled quote
                        s.append(ch)
                        j += 2
                        continue
                    break
                s.append(body[j])
                j += 1
            toks.append(("str", "".join(s)))
            i = j + 1
        elif ch in ", \t\n\r":
            i += 1
        elif ch == "=":
            toks.append(("eq", "="))
            i += 1
        else:
            j = i
            depth = 0
            while j < n and (depth > 0 or body[j] not in ", \t\n\r='\""):
                if body[j] == "(":
                    depth += 1
                elif body[j] == ")":
                    depth -= 1
                j += 1
            tok = body[i:j]
            if j < n and body[j] in "'\"" and tok.endswith("*"):  # repeat count of a string
                toks.append(("rep", int(tok[:-1])))
            else:
                toks.append(("word", tok))
            i = j
    return toks


def _value(tok, typ, name):
    kind, v = tok
    if typ == "s":
        if kind != "str":
            raise ValueError(f"{name}: expected a quoted string, got {v!r}")
        return v
    if kind != "word":
        raise ValueError(f"{name}: expected a {'number' if typ in 'ir' else 'logical'}, got {v!r}")
    if typ == "l":
        w = v.strip(".").lower()
        if w and w[0] in "tf":
            return w[0] == "t"
        raise ValueError(f"{name}: bad logical {v!r}")
    w = v.lower().replace("d", "e")
    return int(w) if typ == "i" else float(w)


def _indices(sub, bounds, name):
    """Linear (column-major) offsets of an array section such as (0:1,1:3,2)."""
    parts = [p.strip() for p in sub.split(",")]
    if len(parts) != len(bounds):
        raise ValueError(f"{name}: {len(parts)} subscripts for a rank-{len(bounds)} array")
    ranges = []
    for p, (lo, hi) in zip(parts, bounds):
        if ":" in p:
            a, b = p.split(":")
            a = int(a) if a.strip() else lo
            b = int(b) if b.strip() else hi
        else:
            a = b = int(p)
        if a < lo or b > hi or a > b:
            raise ValueError(f"{name}: subscript {p} outside {lo}:{hi}")
        ranges.append(range(a - lo, b - lo + 1))
    offs = [0]
    stride = 1
    for r, (lo, hi) in zip(ranges, bounds):
        offs = [o + stride * k for k in r for o in offs]
        stride *= hi - lo + 1
    return offs


def parse(text, vof=False):
    """The values of the reference's namelist groups in `text`, on top of the
    defaults; arrays are flat lists in column-major order."""
    vals = defaults(vof)
    text = _strip_comments(text)
    for m in re.finditer(r"&(\w+)(.*?)(?:^\s*/|/\s*$)", text, re.S | re.M):
        group = m.group(1).lower()
        if group not in GROUPS:
            continue  # e.g. &cudecomp: settings of the reference's GPU decomposition
        toks = _tokens(m.group(2))
        k = 0
        while k < len(toks):
            kind, w = toks[k]
            if kind != "word" or k + 1 >= len(toks) or toks[k + 1][0] != "eq":
                raise ValueError(f"&{group}: expected 'name =' at {w!r}")
            mm = re.fullmatch(r"(\w+)\s*(?:\((.*)\))?", w)
            name, sub = mm.group(1).lower(), mm.group(2)
            if name not in DECL or DECL[name][0] != group:
                raise ValueError(f"&{group}: unknown variable {name}")
            _, typ, bounds = DECL[name]
            k += 2
            items = []
            while k < len(toks) and not (k + 1 < len(toks) and toks[k + 1][0] == "eq" and toks[k][0] == "word"):
                kind2, v2 = toks[k]
                if kind2 == "rep":
                    items += [_value(toks[k + 1], typ, name)] * v2
                    k += 2
                    continue
                if kind2 == "word" and "*" in v2 and not v2.startswith("*"):
                    cnt, rest = v2.split("*", 1)
                    items += [_value(("word", rest), typ, name)] * int(cnt)
                else:
                    items.append(_value(toks[k], typ, name))
                k += 1
            if not bounds:
                if len(items) != 1:
                    raise ValueError(f"{name}: one value expected, got {len(items)}")
                vals[name] = items[0]
                continue
            offs = _indices(sub, bounds, name) if sub else list(range(_size(bounds)))
            if len(items) > len(offs):
                raise ValueError(f"{name}: {len(items)} values for {len(offs)} elements")
            arr = list(vals[name])
            for o, v in zip(offs, items):
                arr[o] = v
            vals[name] = arr
    return vals


def read_spheres(text):
    """The (x, y, z, r) records of a spheres.in file (list-directed, one per line)."""
    out = []
    for line in text.splitlines():
        f = line.replace(",", " ").split()
        if not f:
            continue
        out.append(tuple(float(x.lower().replace("d", "e")) for x in f[:4]))
    return out


INIPSI_ALIASES = {"dis3": "drp3", "dis2": "drp2", "dis1": "drp1"}  # src/two_fluid.f90:290-306


def case_from_input(text, spheres_text=None, vof=False):
    """A case dict of fizzy_amd.cases from an input.nml text (and the
    spheres.in text where inipsi needs one). vof selects THINC/QQ, the
    reference's build flag _INTERFACE_CAPTURING_VOF."""
    v = parse(text, vof)
    if v["inivel"] != "zer":
        raise ValueError(f"inivel '{v['inivel']}': only a fluid at rest ('zer') is supported")
    if v["is_wallturb"]:
        raise ValueError("is_wallturb: not supported")
    inipsi = INIPSI_ALIASES.get(v["inipsi"], v["inipsi"])
    spheres = read_spheres(spheres_text) if spheres_text else []
    keys = ["ng", "l", "gtype", "gr", "cfl", "dtmax", "dt_f", "is_solve_ns", "is_track_interface",
            "is_forced_hit", "icheck", "cbcvel", "bcvel", "cbcpre", "bcpre", "cbcpsi", "bcpsi", "cbcnor", "bcnor",
            "bforce", "gacc", "sigma", "rho12", "mu12", "beta12", "psi_thickness_factor", "nstep", "time_max",
            "tw_max", "stop_type"]
    case = {k: v[k] for k in keys}
    case.update(inipsi=inipsi, spheres=spheres, vof=bool(vof))
    return case
