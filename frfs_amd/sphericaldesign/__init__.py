"""Symmetric spherical designs (R. S. Womersley's point sets, shipped as data
files in ``symmetric/`` with the file naming of the reference,
``<name>-ss<degree>-m<npts>.txt``).  Loader semantics follow
frfs/sphericaldesign/__init__.py:45-107: a stored rule is selected by name
and exact point count, the highest degree winning; no match raises
``ValueError('No suitable spherical quadrature rule found')``."""
import os
import re

import numpy as np

_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'symmetric')


class SphericalRule:
    def __init__(self, pts, wts=None):
        self.pts = pts
        self.wts = np.zeros(0) if wts is None else wts


def _parse(text):
    pts, wts = [], []
    for line in text.splitlines():
        line = re.sub(r'[(),\[\]]', ' ', line).strip()
        if not line:
            continue
        vals = [float(v) for v in line.split()]
        if len(vals) == 3:
            pts.append(vals)
        elif len(vals) == 4:
            pts.append(vals[:3])
            wts.append(vals[3])
        else:
            raise ValueError('Invalid points in quadrature rule')
    if wts and len(wts) != len(pts):
        raise ValueError('Invalid number of weights')
    return SphericalRule(np.array(pts, dtype=np.float64), np.array(wts, dtype=np.float64))


def stored_rules():
    out = []
    for fn in sorted(os.listdir(_DIR)):
        m = re.match(r'([a-zA-Z0-9\-~+]+)-ss(\d+)(?:-m(\d+))?\.txt$', fn)
        if m:
            out.append((fn, m.group(1), int(m.group(2)), int(m.group(3) or -1)))
    return out


def get_sphquadrule(sstype, rule=None, npts=None, qdeg=None, flags=None):
    if sstype != 'symmetric':
        raise ValueError('Only symmetric spherical designs are available')
    if rule and not re.match(r'[a-zA-Z0-9\-~+]+$', rule):
        r = _parse(rule)
        if npts and npts != len(r.pts):
            raise ValueError('Invalid number of points in provided rule')
        if qdeg and not len(r.wts):
            raise ValueError('Provided rule has no quadrature weights')
        return r
    if not npts and not qdeg:
        raise ValueError('Must specify either npts or qdeg')
    best = None
    for fn, rname, rqdeg, rnpts in stored_rules():
        if ((not rule or rule == rname) and (not npts or npts == rnpts)
                and (not qdeg or qdeg <= rqdeg)):
            if (best is None or (npts and rqdeg > best[2])
                    or (qdeg and rnpts < best[1])):
                best = (fn, rnpts, rqdeg)
    if best is None:
        raise ValueError('No suitable spherical quadrature rule found')
    with open(os.path.join(_DIR, best[0])) as fh:
        return _parse(fh.read())
