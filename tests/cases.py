"""Shared test cases: the operator configurations, built both for the oracle
and for the product plugin from one INI description."""
import numpy as np

from frfs_amd.inifile import Inifile

ARGON = dict(T0=273.0, H0=1e-3, rho0=0.00011140244750324529, molarMass0=39.96e-3)


def vhs_cfg(Nv=16, Nrho=8, M=6, dev=5.0, omega=0.5, precision='double', project=0,
            model='vhs-gll', max_chunk=0, fused=1, pack=1):
    return Inifile.from_dict({
        'non-dim': ARGON,
        'constants': {'Nv': Nv, 'NvBatchSize': Nv ** 3, 'Nrho': Nrho},
        'velocity-mesh': {'dev': dev, 'cmax': 0, 'Tmax': 273},
        'spherical-design-rule': {'ssrule': 'womersley', 'M': M},
        'scattering-model': {'type': model, 'omega': omega, 'dRef': 4.17e-10, 'Tref': 273,
                             'precision': precision, 'conservation-projection': project,
                             'max-chunk': max_chunk, 'fused': fused, 'pair-pack': pack},
    })


def vss_cfg(Nv=16, Nrho=8, M=12, dev=5.0, masses=(39.96e-3,), alpha=None, omega=None,
            dref=None, Mpre=32, precision='double', project=0, max_chunk=0, fused=1, lanes=1, pack=1,
            hpair=1, overlap=0):
    ns = len(masses)
    sm = {'type': 'vss-gll', 'ssrulepre': 'womersley', 'Mpre': Mpre,
          'precision': precision, 'conservation-projection': project, 'max-chunk': max_chunk,
          'fused': fused, 'pencil-lanes': lanes, 'pair-pack': pack, 'group-pairs': hpair,
          'overlap': overlap}
    for p in range(ns):
        for q in range(ns):
            pq = f'{p}{q}'
            sm['alpha' + pq] = (alpha or {}).get(pq, 1.4)
            sm['omega' + pq] = (omega or {}).get(pq, 0.74)
            sm['dRef' + pq] = (dref or {}).get(pq, 4.17e-10)
            sm['Tref' + pq] = 273
    vmesh = {'dev': dev, 'cmax': 0, 'Tmax': 273}
    for p, m in enumerate(masses):
        vmesh['molarMass' + str(p + 1)] = m
    return Inifile.from_dict({
        'non-dim': ARGON,
        'constants': {'Nv': Nv, 'NvBatchSize': Nv ** 3, 'Nrho': Nrho},
        'velocity-mesh': vmesh,
        'spherical-design-rule': {'ssrule': 'womersley', 'M': M},
        'scattering-model': sm,
    })


def oracle_vhs(cfg):
    from oracle import fastspectral as fsp
    g = cfg.getfloat
    vm = fsp.velocity_mesh(cfg.getint('constants', 'Nv'), cfg.getint('constants', 'Nrho'),
                           cfg.getint('spherical-design-rule', 'M'), dev=g('velocity-mesh', 'dev'),
                           **ARGON)
    prm = fsp.vhs_parameters(vm, g('scattering-model', 'omega'), g('scattering-model', 'dRef'),
                             g('scattering-model', 'Tref'))
    tab = fsp.vhs_tables(vm, prm['gamma'])
    return vm, prm, tab


def oracle_vss(cfg, ns):
    from oracle import fastspectral as fsp
    g = cfg.getfloat
    masses = [g('velocity-mesh', 'molarMass' + str(p + 1)) for p in range(ns)]
    vm = fsp.velocity_mesh(cfg.getint('constants', 'Nv'), cfg.getint('constants', 'Nrho'),
                           cfg.getint('spherical-design-rule', 'M'), dev=g('velocity-mesh', 'dev'),
                           masses=masses, **ARGON)
    d = lambda key: {f'{p}{q}': g('scattering-model', key + f'{p}{q}')
                     for p in range(ns) for q in range(ns)}
    prm = fsp.vss_parameters(vm, d('alpha'), d('omega'), d('dRef'), d('Tref'),
                             cfg.get('scattering-model', 'ssrulepre'),
                             cfg.getint('scattering-model', 'Mpre'))
    tab = fsp.vss_tables(vm, prm)
    return vm, prm, tab


def rel_l2(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / np.linalg.norm(np.ravel(b)))
