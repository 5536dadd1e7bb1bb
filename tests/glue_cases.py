"""Synthetic physics states for the tests of the GCM-side code (oracle/glue_oracle.py and the
library's glue kernels): a hydrostatic column per GCM column, winds, water, surface fluxes."""
import numpy as np


def physics_state(ncol, pcols, pver, seed=0):
    rng = np.random.default_rng(seed)
    F = lambda *s: np.zeros(s, order="F")
    st = {k: F(pcols, pver) for k in ("zm", "t", "qv", "qcl", "qci", "pmid", "pdel", "u", "v")}
    st.update({k: F(pcols, pver + 1) for k in ("zi", "pint")})
    st.update({k: F(pcols) for k in ("ps", "phis")})
    for i in range(pcols):
        ps = 1.0e5 + rng.uniform(-2e3, 2e3)
        eta = np.linspace(0.002, 1.0, pver + 1) ** 1.4
        pint = eta * ps
        pmid = 0.5 * (pint[1:] + pint[:-1])
        t = 300.0 - 70.0 * (1.0 - pmid / ps) + rng.normal(0, 0.5, pver)
        zi = np.zeros(pver + 1)  # interface heights, hydrostatic from the surface up
        for k in range(pver - 1, -1, -1):
            zi[k] = zi[k + 1] + 287.0 * t[k] / 9.80616 * np.log(pint[k + 1] / pint[k])
        st["pint"][i], st["pmid"][i], st["pdel"][i] = pint, pmid, np.diff(pint)
        st["t"][i], st["zi"][i] = t, zi
        st["zm"][i] = 0.5 * (zi[1:] + zi[:-1])
        st["qv"][i] = 0.018 * (pmid / ps) ** 3 * rng.uniform(0.5, 1.0, pver)
        st["qcl"][i] = np.where(rng.uniform(size=pver) < 0.3, rng.uniform(0, 1e-4, pver), 0.0)
        st["qci"][i] = np.where(rng.uniform(size=pver) < 0.3, rng.uniform(0, 5e-5, pver), 0.0)
        st["u"][i] = rng.normal(5, 8, pver)
        st["v"][i] = rng.normal(0, 5, pver)
        st["ps"][i], st["phis"][i] = ps, rng.uniform(0, 3000)
    return st


def surface(pcols, seed=1):
    rng = np.random.default_rng(seed)
    return {"ocnfrac": rng.uniform(0, 1, pcols), "wsx": rng.normal(0, 0.1, pcols),
            "wsy": rng.normal(0, 0.1, pcols), "shf": rng.uniform(0, 40, pcols),
            "lhf": rng.uniform(0, 200, pcols), "cflx": rng.uniform(0, 200, pcols) / 2.501e6}
