"""Named solver workloads (cfg objects for SimSetup) with their stated sizes.

cold_cone    cold-gas conical nozzle, 10 mm throat, area ratio 4, 10 bar / 300 K,
             ambient 1 bar / 300 K, 256 x 96, Sutherland viscosity without eddy
             viscosity, 2000 steps at CFL 0.5
hot_bell     methalox Rao bell, 20 mm throat, area ratio 25, gamma 1.2, R 380,
             70 bar / 3500 K, ambient 0.3 bar / 250 K, 768 x 256, full viscous
             model (Smagorinsky + mixing length + Ducros), Werner-Wengle wall at
             800 K, Zukoski separation, afterburning; 2 ms simulated
kerolox_injector  impinging-doublet kerolox injector, 1.5 mm orifices, 10 bar drop,
             mixture ratio 2.6, 768 x 256, 65,536 parcels; 1.5 ms simulated
Choices the configuration text leaves open follow the reference's presets
(chamber length and contraction ratio of the 'coldgas' / 'diamonds' /
'injImpinge' presets, physics.js:246-293) and its plume-length rule for the
domain (fitDomain, app.js:274-282).
"""
import ctypes as C
import math

from . import abi
from ._lib import lib
from .sim_setup import default_config

RU = 8314.462618


def theory(cfg, pa):
    out = (C.c_double * 14)()
    lib().nz_nozzle_theory(cfg.Pc, cfg.Tc, cfg.gamma, cfg.Mw, cfg.Dt, cfg.De, pa, out)
    names = ["At", "Ae", "eps", "Me", "pe", "Te", "ue", "cstar", "mdot", "F", "Fvac", "Isp", "c0", "uLim"]
    return dict(zip(names, list(out)))


def fit_domain(cfg, pa):
    """fitDomain (app.js:274-282): nozzle length plus a plume scaled by exit diameter."""
    from .sim_setup import SimSetup
    probe = default_config()
    C.memmove(C.byref(probe), C.byref(cfg), C.sizeof(abi.NzConfig))
    probe.Nx = 64
    s = SimSetup(probe)
    t = theory(cfg, pa)
    ratio = t["pe"] / pa
    Re = s.info.Re
    plume = 2 * Re * max(6, min(16, 6 + 2.5 * math.log10(max(ratio, 1)) * 3))
    return round(min(8, max(0.1, (s.info.xe + plume) * 1.05)) * 20) / 20


def cold_cone(Nx=256, **over):
    cfg = default_config(prop="coldgas", Tc=300, gamma=1.4, Mw=28.0, qAb=0, sAir=1, Dt=0.020, De=0.040, type="cone",
                         coneAngle=15, convAngle=30, CR=6.25, chamberLen=1.4, Pc=10e5, domainL=0.3, Nx=Nx,
                         aspect=96 / 256, cfl=0.5, ambP=1e5, ambT=300, turbulence=0, Tw=800)
    return apply(cfg, over)


def hot_bell(Nx=768, **over):
    cfg = default_config(prop="methalox", Tc=3500, gamma=1.2, Mw=RU / 380, qAb=3.0, sAir=0.8, Dt=0.040, De=0.200,
                         type="bell", bellPct=80, CR=5, chamberLen=1.5, Pc=70e5, Nx=Nx, aspect=256 / 768,
                         ambP=0.3e5, ambT=250, Tw=800, viscous=1, afterburn=1)
    cfg.domainL = fit_domain(cfg, 0.3e5)
    return apply(cfg, over)


def kerolox_injector(Nx=768, **over):
    cfg = default_config(prop="kerolox", Tc=3670, gamma=1.22, Mw=22.0, qAb=6.8, sAir=1.8, Pc=70e5, Dt=0.030,
                         De=0.085, type="bell", bellPct=80, CR=5, chamberLen=6, domainL=0.5, Nx=Nx, aspect=256 / 768,
                         injMode=1, injType="impinge", injRings=4, injDP=10 / 70 * 100, injD=1.5, injAngle=30, OF=2.6)
    return apply(cfg, over)


def apply(cfg, over):
    from .sim_setup import apply_config
    return apply_config(cfg, **over)


# simulated horizons of the workloads (s)
HORIZON = {"cold_cone": None, "hot_bell": 2e-3, "kerolox_injector": 1.5e-3}
STEPS = {"cold_cone": 2000}
