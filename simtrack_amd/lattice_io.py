"""Lattice readers: a MAD-X TWISS table and a Bmad element table into a Line.

Both follow the reference's readers element for element, including their
quirks, so that a file gives the same Line (types, parameters, order) as
Read_MADX_Lattice (simtrack.cpp:13591-13796) and Read_BMAD_Lattice
(cpp:14054-14237) build.  Numbers are parsed as decimal strings, which
rounds exactly as the reference's istringstream >> double does.

A file may be gzip-compressed (name ending in .gz).
"""
import gzip

from . import line as L


def _open(path):
    if str(path).endswith(".gz"):
        return gzip.open(path, "rt")
    return open(path)


class LatticeError(ValueError):
    pass


# MAD-X: columns NAME KEYWORD S L ANGLE E1 E2 TILT K0L K0SL ... K10L K10SL KS
_MADX_NUM = 6 + 22 + 1


def _madx_element(name, kind, l, angle, e1, e2, tilt, knl, knsl, ks):
    """One row of Read_MADX_Lattice (cpp:13634-13790)."""
    if kind == "DRIFT":
        return L.DRIFT(name, l)
    if kind == "SBEND":
        if knl[0] != 0. and knl[2] == 0.:                     # cpp:13657
            e = L.GSBEND(name, l, angle, knl[0], e1, e2)
        elif knl[1] == 0. and knl[2] == 0.:                   # cpp:13662
            e = L.SBEND(name, l, angle, e1, e2)
        else:                                                 # cpp:13667
            e = L.GMULT(name, l, angle, e1, e2, knl, knsl)
        e.DPSI = tilt
        return e
    if kind == "RBEND":                                       # cpp:13674-13684
        if knl[1] == 0. and knl[2] == 0.:
            return L.SBEND(name, l, angle, angle / 2, angle / 2)
        return L.GMULT(name, l, angle, e1, e2, knl, knsl)
    if kind in ("QUADRUPOLE", "SEXTUPOLE", "OCTUPOLE"):
        cls, order = {"QUADRUPOLE": (L.QUAD, 1), "SEXTUPOLE": (L.SEXT, 2), "OCTUPOLE": (L.OCT, 3)}[kind]
        e = cls(name, l, knl[order], knsl[order])
        e.DPSI = tilt
        return e
    # the reader leaves hkick = vkick = 0 (their columns are not read, cpp:13650)
    if kind == "HKICKER":
        return L.HKICKER(name, l, 0.)
    if kind == "VKICKER":
        return L.VKICKER(name, l, 0.)
    if kind == "TKICKER":
        return L.KICKER(name, l, 0., 0.)
    if kind == "MULTIPOLE":                                   # cpp:13745-13749
        e = L.MULT(name, l, knl, knsl)
        e.DPSI = tilt
        return e
    if kind == "VMONITOR":
        return L.VBPM(name, l)
    if kind == "MONITOR":
        return L.BPM(name, l)
    if kind == "MARKER":
        return L.MARKER(name, l)
    if kind == "SOLENOID":                                    # cpp:13762-13771
        if ks != 0. and l != 0.:
            return L.SOLEN(name, l, ks)
        return L.DRIFT(name, l)
    if kind == "RFCAVITY":
        return L.RFCAV(name, l, 0., 0., 0.)
    if kind == "BEAMBEAM":                                    # cpp:13776-13783
        if l != 0.:
            raise LatticeError("Beam-Beam element should have zero length!")
        return L.BEAMBEAM(name, 6, 0., 1, 0.4545, 11, 2.5e-06, 2.5e-06, 0.53, 0.0, 0.53, 0.0)
    if kind == "ELENS":                                       # cpp:13784-13787 (no current)
        return L.ELENS(name, l, 0.0, 1, 0.14, 4, .31e-03, .31e-03)
    return L.DRIFT(name, l)                                   # every other keyword, cpp:13788


def read_madx_twiss(path):
    """Read_MADX_Lattice: a MAD-X TWISS table (columns as above).  Returns
    the Line (Update() called); its `circumference` attribute is the S of
    the last element, as the reader sets GP.circumference."""
    with _open(path) as f:
        lines = f.read().split("\n")
    # the header ends two lines after the one holding "KEYWORD" (cpp:13619-13628)
    nskip = None
    for i, ln in enumerate(lines[:100]):
        if "KEYWORD" in ln:
            nskip = i + 2
            break
    if nskip is None:
        raise LatticeError(f"{path}: no KEYWORD header line in the first 100 lines")
    cells = []
    for ln in lines[nskip:]:
        tok = ln.split()
        if not tok:
            continue
        if len(tok) < 2 + _MADX_NUM:
            raise LatticeError(f"{path}: short row {ln[:60]!r}")
        name, kind = tok[0][1:-1], tok[1][1:-1]
        v = [float(t) for t in tok[2:2 + _MADX_NUM]]
        s, l, angle, e1, e2, tilt = v[:6]
        knl = [v[6 + 2 * k] for k in range(11)]
        knsl = [v[7 + 2 * k] for k in range(11)]
        ks = v[28]
        cells.append(_madx_element(name, kind, l, angle, e1, e2, tilt, knl, knsl, ks))
    ln = L.Line(cells)
    ln.Update()
    ln.circumference = ln.Cell[-1].S if ln.Cell else 0.
    return ln


# Bmad: name key s l angle actual_angle e1 e2 tilt x_offset z_offset x_pitch
#       k0l k0sl ... k10l k10sl ks hkick vkick
_BMAD_NUM = 10 + 22 + 3


def _bmad_element(name, kind, l, angle, e1, e2, xoff, zoff, xpitch, knl, knsl, ks, hkick, vkick):
    """One row of Read_BMAD_Lattice (cpp:14104-14230)."""
    if kind == "Drift":
        return L.DRIFT(name, l)
    if kind == "Sbend":
        if abs(angle) < 1.0e-9:                               # cpp:14109
            return L.MULT(name, l, knl, knsl)
        if abs(angle - knl[0]) > 1.0e-6:                      # cpp:14113
            return L.GSBEND(name, l, angle, knl[0], e1, e2)
        knl = [0.] + list(knl[1:])                            # cpp:14118
        if knl[2] == 0. and knl[1] == 0.:
            return L.SBEND(name, l, angle, e1, e2)
        return L.SBENDMULT(name, l, angle, e1, e2, knl, knsl)
    if kind == "RBEND":                                       # cpp:14129-14137
        if knl[1] == 0. and knl[2] == 0.:
            return L.SBEND(name, l, angle, angle / 2., angle / 2.)
        return L.SBENDMULT(name, l, angle, angle / 2., angle / 2., knl, knsl)
    if kind == "Quadrupole":
        return L.QUAD(name, l, knl[1], knsl[1])
    if kind == "Sextupole":
        return L.SEXT(name, l, knl[2], knsl[2])
    if kind == "OCTUPOLE":
        return L.OCT(name, l, knl[3], knsl[3])
    if kind == "HKICKER":
        return L.HKICKER(name, l, hkick)
    if kind == "VKICKER":
        return L.VKICKER(name, l, vkick)
    if kind == "Kicker":
        return L.KICKER(name, l, hkick, vkick)
    if kind == "Multipole":                                   # cpp:14166-14180
        if knl[0] == 0:
            return L.MULT(name, l, knl, knsl)
        return L.SBEND(name + "_tb", 1.0e-6, knl[0], 0, 0)
    if kind == "HMONITOR":
        return L.HBPM(name, l)
    if kind == "VMONITOR":
        return L.VBPM(name, l)
    if kind == "Monitor":
        return L.BPM(name, l)
    if kind == "MARKER":
        return L.MARKER(name, l)
    if kind == "Solenoid":                                    # cpp:14197-14206
        if l != 0.:
            return L.SOLEN(name, l, ks)
        return L.DRIFT(name, l)
    if kind == "RFcavity":
        return L.RFCAV(name, l, 0., 0., 0.)
    if kind == "BEAMBEAM":
        if l != 0.:
            raise LatticeError("Beam-Beam element should have zero length!")
        return L.BEAMBEAM(name, 6, 0., 1, 0.4545, 11, 2.5e-06, 2.5e-06, 0.53, 0.0, 0.53, 0.0)
    if kind == "ELENS":                                       # cpp:14218-14221 (no current)
        return L.ELENS(name, l, 0.0, 1, 0.14, 4, .31e-03, .31e-03)
    if kind == "Patch":                                       # cpp:14222-14225
        return L.PATCH(name, xoff, 0.0, zoff, -xpitch)
    return L.DRIFT(name, l)                                   # cpp:14226-14230


def read_bmad_table(path):
    """Read_BMAD_Lattice: a Bmad element table.  Only the first line is
    skipped (cpp:14085), so the column-name line and the units line are read
    as rows too: their second word is no known key and their numbers do not
    parse, which makes each a DRIFT whose length the reference takes from an
    unset variable; here that length is 0.  Keys are matched with the
    reference's capitalisation (so "Hkicker", "Marker" and "Instrument"
    become drifts)."""
    with _open(path) as f:
        lines = f.read().split("\n")
    cells = []
    for ln in lines[1:]:
        tok = ln.split()
        if not tok:
            continue
        name, kind = tok[0], tok[1] if len(tok) > 1 else ""
        try:
            v = [float(t) for t in tok[2:2 + _BMAD_NUM]]
            ok = len(v) == _BMAD_NUM
        except ValueError:
            ok = False
        if not ok:
            cells.append(L.DRIFT(name, 0.))
            continue
        s, l, angle, actual, e1, e2, tilt, xoff, zoff, xpitch = v[:10]
        knl = [v[10 + 2 * k] for k in range(11)]
        knsl = [v[11 + 2 * k] for k in range(11)]
        ks, hkick, vkick = v[32], v[33], v[34]
        cells.append(_bmad_element(name, kind, l, angle, e1, e2, xoff, zoff, xpitch, knl, knsl, ks, hkick, vkick))
    ln = L.Line(cells)
    ln.Update()
    ln.circumference = ln.Cell[-1].S if ln.Cell else 0.
    return ln
