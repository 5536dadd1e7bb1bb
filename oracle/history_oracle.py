"""CPU restatement of the CRM history fields (TEST INFRASTRUCTURE ONLY): the outfld calls of
crm_physics_tend (components/cam/src/physics/cam/crm_physics.F90:1201-1461, sam1mom) as
time means over several calls.  Only tests/ import this module.

Arrays are numpy arrays in the Fortran layouts of their sources, indexed [column, ...]
((pcols, nx, ny, nz), (pcols, pver), (pcols)).  Each field's value per call follows the
reference expression in the reference's evaluation order; time means are the sequential sum of
the calls' values times 1/count, as the device history computes them.
"""
import numpy as np

CPAIR = 1.00464e3                     # shr_const_mod.F90:47 via physconst
KG_PER_G = float(np.float32(1.0e-3))  # crm_physics.F90:1406-1409, a 4-byte literal


def incloud(path, cld):
    """crm_physics.F90:1406-1407: in-cloud water path [kg/m2] from a grid-mean path [g/m2]."""
    return path * KG_PER_G / np.maximum(0.01, cld)


def column_path(path):
    """crm_physics.F90:1408-1409: sum over levels, top first, of path * 1e-3 per column."""
    tot = np.zeros(path.shape[0])
    for k in range(path.shape[1]):
        tot = tot + path[:, k] * KG_PER_G
    return tot


def fields(state=None, out=None, qrad=None, pmid=None, pdel=None, ptend=None, qrs=None, qrl=None):
    """{history name: value of one call} for every field whose inputs are given (dicts of
    crm_state / crm_output members; ptend a dict s, qv, qcl, qci)."""
    f = {}
    if qrad is not None:
        f["CRM_QRAD"] = qrad                                            # :1201
    if pmid is not None:
        f["PRES"] = pmid                                                # :1211
    if pdel is not None:
        f["DPRES"] = pdel                                               # :1212
    if state is not None:                                               # :1214-1217
        f.update({"CRM_U": state["u_wind"], "CRM_V": state["v_wind"], "CRM_W": state["w_wind"],
                  "CRM_T": state["temperature"]})
    if state is not None and out is not None:
        f["CRM_QV"] = state["qt"] - out["qcl"] - out["qci"]             # :1220
    if out is not None:
        f.update({"CRM_QC": out["qcl"], "CRM_QI": out["qci"], "CRM_QPC": out["qpl"],  # :1224-1230
                  "CRM_QPI": out["qpi"], "CRM_PREC": out["prec_crm"], "CRM_TK": out["tk"],
                  "CRM_TKH": out["tkh"]})
    if qrl is not None:
        f["SPQRL"] = qrl / CPAIR                                        # :1317
    if qrs is not None:
        f["SPQRS"] = qrs / CPAIR                                        # :1318
    if ptend is not None:                                               # :1338-1340
        f.update({"SPDQ": ptend["qv"], "SPDQC": ptend["qcl"], "SPDQI": ptend["qci"]})
    if ptend is not None and qrs is not None and qrl is not None:
        f["SPDT"] = (ptend["s"] - qrs - qrl) / CPAIR                    # :1333,1341
    if out is not None:
        copy = {"SPMC": "mctot", "SPMCUP": "mcup", "SPMCDN": "mcdn", "SPMCUUP": "mcuup",   # :1342-1378
                "SPMCUDN": "mcudn", "SPQC": "qc_mean", "SPQI": "qi_mean", "SPQS": "qs_mean",
                "SPQG": "qg_mean", "SPQR": "qr_mean", "SPQTFLX": "flux_qt", "SPUFLX": "flux_u",
                "SPVFLX": "flux_v", "SPTKE": "tkez", "SPTKES": "tkesgsz", "SPTK": "tkz",
                "SPQTFLXS": "fluxsgs_qt", "SPQPFLX": "flux_qp", "SPPFLX": "precflux",
                "SPQTLS": "qt_ls", "SPQTTR": "qt_trans", "SPQPTR": "qp_trans", "SPQPEVP": "qp_evp",
                "SPQPFALL": "qp_fall", "SPQPSRC": "qp_src", "SPTLS": "t_ls",
                "CLOUD": "cld", "CLDTOT": "cltot", "CLDHGH": "clhgh", "CLDMED": "clmed",  # :1390-1397
                "CLDLOW": "cllow", "CLOUDTOP": "cldtop", "TIMINGF": "timing_factor",
                "MU_CRM": "mu_crm", "MD_CRM": "md_crm", "EU_CRM": "eu_crm",          # :1457-1461
                "DU_CRM": "du_crm", "ED_CRM": "ed_crm"}
        f.update({k: out[v] for k, v in copy.items()})
        gi, gl, cld = out["gicewp"], out["gliqwp"], out["cld"]
        f["GCLDLWP"] = gi + gl                                          # :1413,1416
        ti, tl = column_path(gi), column_path(gl)
        f["TGCLDCWP"] = ti + tl                                         # :1412,1417
        f["TGCLDLWP"] = tl                                              # :1418
        f["TGCLDIWP"] = ti                                              # :1419
        f["ICLDTWP"] = incloud(gi, cld) + incloud(gl, cld)              # :1414,1420
        f["ICLDIWP"] = incloud(gi, cld)                                 # :1421
    return f


class History:
    """Time means over calls of the fields of the chunk's first ncol columns."""

    def __init__(self):
        self.sums, self.count = {}, {}

    def accumulate(self, ncol, **inputs):
        for name, v in fields(**inputs).items():
            v = np.asarray(v, dtype=np.float64)[:ncol]
            if name in self.sums:
                self.sums[name] = self.sums[name] + v
            else:
                self.sums[name] = 0.0 + v
            self.count[name] = self.count.get(name, 0) + 1

    def mean(self, name):
        return self.sums[name] * (1.0 / self.count[name]), self.count[name]
