"""Model construction: mechanism sets, seeded morphologies, cable coefficients,
region parameter maps and the benchmark configurations.

A model here is everything the reference's porting step (``PYS/extractModel.py``)
would produce from NEURON: the tree file contents (:class:`formats.TreeData`),
the parameter-matrix layout (``AllParams``, index ``p*NComps + comp``), the
stimulus segment and area, v_init and the record sites.  Morphologies are
generated instead of exported from NEURON; the discretisation follows NEURON's
compartment conventions (see :func:`cable_coefficients`).
"""
from __future__ import annotations

import dataclasses
import math
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from . import tree as treemod
from .formats import TreeData

# ----------------------------------------------------------------- mechanism sets
# Mechanism order is alphabetical, case-insensitive (PYS/extractModel.py:327);
# parameter order within a mechanism follows the generated signatures.
MECHSETS: Dict[str, dict] = {
    "passive": dict(id=0, mechs=["pas"], params=["g_pas", "e_pas"]),
    "ttpc": dict(id=1, mechs=["Ca_HVA", "Ca_LVAst", "CaDynamics_E2", "Ih", "Im", "K_Pst", "K_Tst",
                               "Nap_Et2", "NaTa_t", "NaTs2_t", "pas", "SK_E2", "SKv3_1"],
                 params=["gCa_HVAbar_Ca_HVA", "gCa_LVAstbar_Ca_LVAst", "gamma_CaDynamics_E2",
                         "decay_CaDynamics_E2", "depth_CaDynamics_E2", "minCai_CaDynamics_E2",
                         "gIhbar_Ih", "ehcn_Ih", "gImbar_Im", "gK_Pstbar_K_Pst", "gK_Tstbar_K_Tst",
                         "gNap_Et2bar_Nap_Et2", "gNaTa_tbar_NaTa_t", "gNaTs2_tbar_NaTs2_t",
                         "g_pas", "e_pas", "gSK_E2bar_SK_E2", "zTau_SK_E2", "gSKv3_1bar_SKv3_1"]),
    "mainen": dict(id=2, mechs=["ca", "cad", "kca", "km", "kv", "na", "pas"],
                   params=["gbar_ca", "cao_ca", "gbar_kca", "caix_kca", "Ra_kca", "Rb_kca",
                           "gbar_km", "tha_km", "qa_km", "Ra_km", "Rb_km",
                           "gbar_kv", "tha_kv", "qa_kv", "Ra_kv", "Rb_kv",
                           "gbar_na", "tha_na", "qa_na", "Ra_na", "Rb_na", "thi1_na", "thi2_na",
                           "qi_na", "thinf_na", "qinf_na", "Rg_na", "Rd_na", "g_pas", "e_pas"]),
}

REGIONS = ("soma", "axon", "basal", "apical")

# Entry value marking free parameter i in a template parameter row
# (create_params_to_all_params, Figures/Figure4_DEAP/extractModel_mappings.py:258-270).
SENTINEL = 12300.0


# ----------------------------------------------------------------- morphology
@dataclasses.dataclass
class Morphology:
    """Compartments in Hines order (parent(s) < s, root 0).  Lengths and
    diameters in um; ``section`` groups compartments into NEURON-like sections
    (the reference's "comps", the columns of the parameter matrix)."""

    parent: np.ndarray
    length: np.ndarray
    diam: np.ndarray
    region: np.ndarray
    section: np.ndarray
    ext: Optional[np.ndarray] = None      # extension compartments of a dispersed branching

    @property
    def N(self) -> int:
        return int(self.parent.size)

    @property
    def nsections(self) -> int:
        return int(self.section.max()) + 1

    def area(self) -> np.ndarray:
        return math.pi * self.diam * self.length

    def distance(self) -> np.ndarray:
        """Path distance (um) from the root to each compartment's centre."""
        d = np.zeros(self.N)
        for s in range(1, self.N):
            p = self.parent[s]
            d[s] = d[p] + 0.5 * self.length[p] + 0.5 * self.length[s]
        return d


EXT_LENGTH = 1e-9      # um, the extension section of runModel.hoc:168
EXT_DIAM = 1.0         # um, runModel.hoc:169


def disperse(m: Morphology) -> Morphology:
    """Branch points with more than two children (the root: more than one)
    get a chain of extension compartments that carries the children at most
    two per compartment (:func:`tree.disperse_branching`), each extension a
    cylinder of 1e-9 um x 1 um (runModel.hoc:168-169) without mechanisms, in
    the host's region and a section of its own.  The cable coefficients are
    then computed from the geometry as for any compartment, so the axial
    path from a host to each child is still half the host plus half the
    child.  A morphology that needs no dispersal is returned unchanged."""
    newpar, src = treemod.disperse_branching(m.parent)
    if np.array_equal(src, np.arange(m.N)):
        return m
    orig = src >= 0
    host = np.where(orig, src, -src - 1)
    nsec = m.nsections
    sec = np.where(orig, m.section[host], 0)
    hosts = np.unique(host[~orig])
    for k, h in enumerate(hosts):
        sec[(~orig) & (host == h)] = nsec + k
    return Morphology(newpar, np.where(orig, m.length[host], EXT_LENGTH), np.where(orig, m.diam[host], EXT_DIAM),
                      m.region[host].copy(), sec.astype(np.int64), ~orig)


def cable_coefficients(m: Morphology, Ra) -> Tuple[np.ndarray, np.ndarray]:
    """Off-diagonal Hines coefficients e, f in Hines-row order (row h = N-1-s).

    Between compartment s and its parent p the axial resistance is the sum of
    the two half-cylinders, r = 1e-2 * Ra * (L_s/2 / A_s + L_p/2 / A_p) MOhm
    (Ra in Ohm cm, lengths um, cross sections A = pi d^2/4 um^2).  As NEURON's
    NODEB/NODEA, f = -100/(r * area_s) is the coupling in the child's row and
    e = -100/(r * area_p) the coupling in the parent's row (S/cm^2; the factor
    100 converts nA/(mV um^2) to mA/(mV cm^2))."""
    N = m.N
    Ra = np.broadcast_to(np.asarray(Ra, dtype=np.float64), (N,))
    area = m.area()
    xs = math.pi * m.diam ** 2 / 4.0
    e = np.zeros(N)
    f = np.zeros(N)
    for s in range(1, N):
        p = m.parent[s]
        r = 1e-2 * (Ra[s] * 0.5 * m.length[s] / xs[s] + Ra[p] * 0.5 * m.length[p] / xs[p])
        h = N - 1 - s
        f[h] = -100.0 / (r * area[s])
        e[h] = -100.0 / (r * area[p])
    return e, f


class _Builder:
    """Accumulates sections, then numbers compartments depth-first."""

    def __init__(self):
        self.secs: List[dict] = []

    def add(self, parent: int, region: int, L: float, d0: float, d1: float, nseg: int) -> int:
        self.secs.append(dict(parent=parent, region=region, L=L, d0=d0, d1=d1, nseg=max(1, nseg), kids=[]))
        k = len(self.secs) - 1
        if parent >= 0:
            self.secs[parent]["kids"].append(k)
        return k

    def ncomp(self) -> int:
        return sum(s["nseg"] for s in self.secs)

    def build(self) -> Morphology:
        par, L, D, R, S = [], [], [], [], []
        stack = [(0, -1)]
        while stack:
            k, pcomp = stack.pop()
            sec = self.secs[k]
            n = sec["nseg"]
            for j in range(n):
                par.append(pcomp if j == 0 else len(par) - 1)
                L.append(sec["L"] / n)
                x = (j + 0.5) / n
                D.append(sec["d0"] + (sec["d1"] - sec["d0"]) * x)
                R.append(sec["region"])
                S.append(k)
            last = len(par) - 1
            for kid in reversed(sec["kids"]):
                stack.append((kid, last))
        return Morphology(np.array(par, np.int64), np.array(L), np.array(D), np.array(R, np.int64),
                          np.array(S, np.int64))


def binary_tree(seed: int = 1000, depth: int = 5) -> Morphology:
    """Seeded binary tree: a soma (root, L = d = 20 um) with one child that is
    the root of a full binary tree of ``depth``+1 levels (63 compartments for
    depth 5), L ~ U[10, 50] um, diam ~ U[0.5, 2] um; 64 compartments in all.
    Every compartment is its own section."""
    rng = np.random.default_rng(seed)
    b = _Builder()
    soma = b.add(-1, 0, 20.0, 20.0, 20.0, 1)
    level = [b.add(soma, 2, rng.uniform(10, 50), *(2 * [rng.uniform(0.5, 2.0)]), 1)]
    for _ in range(depth):
        nxt = []
        for k in level:
            for _c in range(2):
                d = rng.uniform(0.5, 2.0)
                nxt.append(b.add(k, 2, rng.uniform(10, 50), d, d, 1))
        level = nxt
    return b.build()


def pyramidal(seed: int, n_target: int, n_basal: int = 4, seg_len: float = 20.0,
              axon_len: float = 600.0, apical_bias: float = 0.0) -> Morphology:
    """Seeded pyramidal-like morphology with about ``n_target`` compartments.

    The soma is a chain of short pieces so that every compartment has at most
    two children and the root one (the reference's solver schedule requires
    it, create_auxilliary_data_3.py:83-84).  Trees: an axon (initial segment +
    unbranched cable), an apical trunk with Galton-Watson obliques and tuft,
    and ``n_basal`` Galton-Watson basal trees.  Growth is breadth-first over
    all trees with a compartment budget; each section branches into two
    children with a depth-dependent probability, and while the budget is not
    met a random terminal section restarts the process (an apical one with
    probability ``apical_bias``)."""
    rng = np.random.default_rng(seed)
    b = _Builder()
    ntrees = n_basal + 2
    soma_pieces = [b.add(-1, 0, 20.0 / ntrees, 20.0, 20.0, 1)]
    for _ in range(ntrees - 1):
        soma_pieces.append(b.add(soma_pieces[-1], 0, 20.0 / ntrees, 20.0, 20.0, 1))
    hosts = soma_pieces[1:] + [soma_pieces[-1]]
    ais = b.add(hosts[0], 1, 40.0, 1.5, 1.2, 4)
    b.add(ais, 1, axon_len, 1.0, 1.0, max(1, int(round(axon_len / 40.0))))
    frontier = []
    trunk = b.add(hosts[1], 3, 300.0, 4.0, 2.0, int(300 / seg_len))
    frontier.append((trunk, 3, 1, 2.0))
    for i in range(n_basal):
        L = rng.uniform(40, 120)
        k = b.add(hosts[2 + i], 2, L, 1.6, 1.2, max(1, int(round(L / seg_len))))
        frontier.append((k, 2, 1, 1.2))
    terminals = []
    while b.ncomp() < n_target:
        if not frontier:
            if not terminals:
                break
            apical = [i for i, tt in enumerate(terminals) if tt[1] == 3]
            pick = apical[int(rng.integers(len(apical)))] if apical and rng.random() < apical_bias \
                else int(rng.integers(len(terminals)))
            k, region, depth, d = terminals.pop(pick)
            frontier.append((k, region, 1, d))
            continue
        idx = int(rng.integers(len(frontier)))
        k, region, depth, d = frontier.pop(idx)
        pbranch = 0.9 if depth < 3 else max(0.25, 0.85 - 0.1 * depth)
        if rng.random() >= pbranch:
            terminals.append((k, region, depth, d))
            continue
        for _ in range(2):
            room = n_target - b.ncomp()
            if room <= 0:
                break
            L = rng.uniform(60, 200) if region == 3 else rng.uniform(40, 150)
            nseg = min(room, max(1, int(round(L / seg_len))))
            dk = max(0.4, d * 0.75)
            kid = b.add(k, region, L * nseg / max(1, int(round(L / seg_len))), d, dk, nseg)
            frontier.append((kid, region, depth + 1, dk))
    return b.build()


def branchy(seed: int = 1004) -> Morphology:
    """A cell whose branch points have more than two children: a one-compartment
    soma carrying an axon and five dendrites (six children), and a dendrite
    that splits three ways at its end, twice.  Such trees are dispersed
    before scheduling (:func:`disperse`)."""
    rng = np.random.default_rng(seed)
    b = _Builder()
    soma = b.add(-1, 0, 20.0, 16.0, 16.0, 1)
    b.add(soma, 1, 400.0, 1.2, 1.0, 10)
    for i in range(5):
        L = rng.uniform(80, 160)
        k = b.add(soma, 2 + i % 2, L, 1.8, 1.2, max(1, int(round(L / 20))))
        if i in (1, 3):
            for _ in range(3):
                L2 = rng.uniform(40, 100)
                k2 = b.add(k, 2 + i % 2, L2, 1.0, 0.7, max(1, int(round(L2 / 20))))
                if i == 3:
                    for _ in range(3):
                        b.add(k2, 3, 30.0, 0.6, 0.5, 2)
    return b.build()


# ----------------------------------------------------------------- models
@dataclasses.dataclass
class FreeParam:
    name: str
    param: str                    # parameter name of the mechanism set
    regions: Tuple[str, ...]      # regions whose sections it sets
    nominal: float
    lo: float                     # log-uniform sampling range
    hi: float


@dataclasses.dataclass
class Model:
    name: str
    mechset: str
    morph: Morphology
    tree: TreeData
    base: np.ndarray              # [nparams, ncomps] parameter values before free params
    free: List[FreeParam]
    v_init: float
    stim_seg: int
    stim_area: float
    rec_sites: np.ndarray

    @property
    def mechset_id(self) -> int:
        return MECHSETS[self.mechset]["id"]

    @property
    def nparams(self) -> int:
        return len(MECHSETS[self.mechset]["params"])

    @property
    def ncomps(self) -> int:
        return self.tree.ncomps

    def section_regions(self) -> np.ndarray:
        reg = np.zeros(self.ncomps, np.int64)
        reg[self.morph.section] = self.morph.region
        return reg

    def allparams(self, x: np.ndarray) -> np.ndarray:
        """AllParams rows [nsets, nparams*ncomps] for free-parameter values
        x [nsets, nfree] (index p*NComps + comp, proc_add_param_to_hoc_for_opt.py:302-310)."""
        x = np.atleast_2d(np.asarray(x, dtype=np.float64))
        if x.shape[1] != len(self.free):
            raise ValueError(f"expected {len(self.free)} free parameters, got {x.shape[1]}")
        pidx = {p: i for i, p in enumerate(MECHSETS[self.mechset]["params"])}
        sreg = self.section_regions()
        out = np.repeat(self.base.reshape(1, -1), x.shape[0], axis=0)
        for j, fp in enumerate(self.free):
            cols = [pidx[fp.param] * self.ncomps + c for c in range(self.ncomps)
                    if REGIONS[sreg[c]] in fp.regions]
            out[:, cols] = x[:, j:j + 1]
        return out

    def param_template(self) -> np.ndarray:
        """One parameter row in which free parameter i sits at every entry it
        sets as the sentinel 12300 + i (the GA evaluators' template)."""
        return self.allparams(np.array([[SENTINEL + i for i in range(len(self.free))]]))[0]

    def nominal(self) -> np.ndarray:
        return np.array([fp.nominal for fp in self.free])

    def sample(self, n: int, seed: int) -> np.ndarray:
        """Free parameters log-uniform in [lo, hi] (seeded)."""
        rng = np.random.default_rng(seed)
        lo = np.log10([fp.lo for fp in self.free])
        hi = np.log10([fp.hi for fp in self.free])
        return 10 ** rng.uniform(lo, hi, size=(n, len(self.free)))

    def state_count(self) -> float:
        """Mean number of FP64 state variables per compartment (gates plus cai),
        the S-bar of the algorithmic byte count."""
        per = {"passive": {}, "ttpc": {"Ca_HVA": 2, "Ca_LVAst": 2, "CaDynamics_E2": 1, "Ih": 1, "Im": 1,
                                         "K_Pst": 2, "K_Tst": 2, "Nap_Et2": 2, "NaTa_t": 2, "NaTs2_t": 2,
                                         "SK_E2": 1, "SKv3_1": 1},
               "mainen": {"ca": 2, "cad": 1, "kca": 1, "km": 1, "kv": 1, "na": 2}}[self.mechset]
        mechs = MECHSETS[self.mechset]["mechs"]
        tot = 0
        for m, name in enumerate(mechs):
            tot += per.get(name, 0) * int(self.tree.bool_model[m].sum())
        return tot / self.tree.N


def _assemble(name, mechset, morph, placement, base_by_region, free, v_init, Ra, cm_by_region) -> Model:
    mechs = MECHSETS[mechset]["mechs"]
    params = MECHSETS[mechset]["params"]
    morph = disperse(morph)
    N = morph.N
    regnames = [REGIONS[r] for r in morph.region]
    ext = morph.ext if morph.ext is not None else np.zeros(N, bool)
    bool_model = np.zeros((len(mechs), N), np.int64)
    for m, mech in enumerate(mechs):
        for s in range(N):
            if mech in placement[regnames[s]] and not ext[s]:
                bool_model[m, s] = 1
    e, f = cable_coefficients(morph, Ra)
    cm = np.array([cm_by_region[r] for r in regnames])
    ncomps = morph.nsections
    t = treemod.build_tree(morph.parent, e, f, cm, morph.section, bool_model, ncomps)
    base = np.zeros((len(params), ncomps))
    sreg = np.zeros(ncomps, np.int64)
    sreg[morph.section] = morph.region
    for c in range(ncomps):
        vals = base_by_region[REGIONS[sreg[c]]]
        for p, pname in enumerate(params):
            base[p, c] = vals.get(pname, 0.0)
    soma = 0
    return Model(name, mechset, morph, t, base.reshape(-1), free, v_init, stim_seg=soma,
                 stim_area=float(morph.area()[soma]), rec_sites=np.array([soma], np.int64))


def passive_model(seed: int = 1000) -> Model:
    """Passive branched tree (64 compartments), pas everywhere, e_pas = -70 mV,
    Ra = 100 Ohm cm, cm = 1 uF/cm^2; g_pas is the free parameter."""
    m = binary_tree(seed)
    placement = {r: ("pas",) for r in REGIONS}
    base = {r: {"g_pas": 1e-4, "e_pas": -70.0} for r in REGIONS}
    free = [FreeParam("g_pas", "g_pas", REGIONS, 1e-4, 1e-5, 1e-3)]
    return _assemble("passive64", "passive", m, placement, base, free, -70.0, 100.0, {r: 1.0 for r in REGIONS})


_MAINEN_KIN = {"cao_ca": 2.5, "caix_kca": 1.0, "Ra_kca": 0.01, "Rb_kca": 0.02,
               "tha_km": -30.0, "qa_km": 9.0, "Ra_km": 0.001, "Rb_km": 0.001,
               "tha_kv": 25.0, "qa_kv": 9.0, "Ra_kv": 0.02, "Rb_kv": 0.002,
               "tha_na": -35.0, "qa_na": 9.0, "Ra_na": 0.182, "Rb_na": 0.124, "thi1_na": -50.0,
               "thi2_na": -75.0, "qi_na": 5.0, "thinf_na": -65.0, "qinf_na": 6.2, "Rg_na": 0.0091,
               "Rd_na": 0.024, "e_pas": -70.0}


def active_model(seed: int = 1001, n_target: int = 500) -> Model:
    """Active pyramidal-like cell with the Mainen & Sejnowski channel set
    (na, kv, km, ca, kca, cad, pas; kinetics as the reference's FigureS4 model)
    in 4 regions, and 8 free conductances."""
    m = pyramidal(seed, n_target)
    dend = ("na", "km", "kca", "ca", "cad", "pas")
    placement = {"soma": ("na", "kv", "km", "kca", "ca", "cad", "pas"), "axon": ("na", "kv", "pas"),
                 "basal": dend, "apical": dend}
    nom = dict(gna=40.0, gna_ax=3000.0, gkv=200.0, gkv_ax=1000.0, gkm=0.1, gca=0.3, gkca=3.0, gpas=1.0 / 30000)
    base = {}
    for r in REGIONS:
        v = dict(_MAINEN_KIN)
        v.update(gbar_na=nom["gna_ax"] if r == "axon" else nom["gna"],
                 gbar_kv=nom["gkv_ax"] if r == "axon" else (nom["gkv"] if r == "soma" else 0.0),
                 gbar_km=nom["gkm"], gbar_ca=nom["gca"], gbar_kca=nom["gkca"], g_pas=nom["gpas"])
        base[r] = v
    SD = ("soma", "basal", "apical")
    free = [
        FreeParam("gna_somadend", "gbar_na", SD, nom["gna"], nom["gna"] / 10, nom["gna"] * 10),
        FreeParam("gna_axon", "gbar_na", ("axon",), nom["gna_ax"], nom["gna_ax"] / 10, nom["gna_ax"] * 10),
        FreeParam("gkv_soma", "gbar_kv", ("soma",), nom["gkv"], nom["gkv"] / 10, nom["gkv"] * 10),
        FreeParam("gkv_axon", "gbar_kv", ("axon",), nom["gkv_ax"], nom["gkv_ax"] / 10, nom["gkv_ax"] * 10),
        FreeParam("gkm", "gbar_km", SD, nom["gkm"], nom["gkm"] / 10, nom["gkm"] * 10),
        FreeParam("gca", "gbar_ca", SD, nom["gca"], nom["gca"] / 10, nom["gca"] * 10),
        FreeParam("gkca", "gbar_kca", SD, nom["gkca"], nom["gkca"] / 10, nom["gkca"] * 10),
        FreeParam("g_pas", "g_pas", REGIONS, nom["gpas"], nom["gpas"] / 10, nom["gpas"] * 10),
    ]
    return _assemble(f"active{m.N}", "mainen", m, placement, base, free, -70.0, 150.0,
                     {r: 0.75 for r in REGIONS})


def branchy_active_model(seed: int = 1004) -> Model:
    """:func:`active_model`'s channels and free conductances on the
    :func:`branchy` morphology (branch points of up to six children)."""
    ref = active_model()
    m = branchy(seed)
    dend = ("na", "km", "kca", "ca", "cad", "pas")
    placement = {"soma": ("na", "kv", "km", "kca", "ca", "cad", "pas"), "axon": ("na", "kv", "pas"),
                 "basal": dend, "apical": dend}
    base = {}
    sreg = ref.section_regions()
    B = ref.base.reshape(ref.nparams, ref.ncomps)
    for r, rn in enumerate(REGIONS):
        c = int(np.nonzero(sreg == r)[0][0])
        base[rn] = {pn: float(B[p, c]) for p, pn in enumerate(MECHSETS["mainen"]["params"])}
    return _assemble(f"branchy{m.N}", "mainen", m, placement, base, ref.free, -70.0, 150.0,
                     {r: 0.75 for r in REGIONS})


def ttpc_model(seed: int = 1002, n_target: int = 1200) -> Model:
    """Thick-tufted-pyramidal-like cell with the BBP TTPC mechanisms placed as
    in the reference's cADpyr232 biophysics (Figures/Figure2_BBP_TTPC_clean/
    biophysics.hoc) and 12 free conductances around its densities."""
    m = pyramidal(seed, n_target, n_basal=5, axon_len=400.0, apical_bias=0.6)
    placement = {
        "basal": ("Ih", "pas"),
        "apical": ("Im", "NaTs2_t", "SKv3_1", "Ih", "pas"),
        "soma": ("Ca_HVA", "SKv3_1", "SK_E2", "Ca_LVAst", "Ih", "NaTs2_t", "CaDynamics_E2", "pas"),
        "axon": ("Ca_HVA", "SKv3_1", "SK_E2", "CaDynamics_E2", "Nap_Et2", "K_Pst", "K_Tst", "Ca_LVAst",
                 "NaTa_t", "pas"),
    }
    common = {"g_pas": 3e-5, "e_pas": -75.0, "ehcn_Ih": -45.0, "zTau_SK_E2": 1.0, "depth_CaDynamics_E2": 0.1,
              "minCai_CaDynamics_E2": 1e-4}
    base = {r: dict(common) for r in REGIONS}
    base["basal"].update(gIhbar_Ih=8e-5)
    base["apical"].update(gNaTs2_tbar_NaTs2_t=0.026145, gSKv3_1bar_SKv3_1=0.004226, gIhbar_Ih=8e-5,
                          gImbar_Im=0.000143)
    base["axon"].update(gNaTa_tbar_NaTa_t=3.137968, gK_Tstbar_K_Tst=0.089259, gamma_CaDynamics_E2=0.002910,
                        gNap_Et2bar_Nap_Et2=0.006827, gSK_E2bar_SK_E2=0.007104, gCa_HVAbar_Ca_HVA=0.000990,
                        gK_Pstbar_K_Pst=0.973538, gSKv3_1bar_SKv3_1=1.021945, decay_CaDynamics_E2=287.198731,
                        gCa_LVAstbar_Ca_LVAst=0.008752)
    base["soma"].update(gamma_CaDynamics_E2=0.000609, gSKv3_1bar_SKv3_1=0.303472, gSK_E2bar_SK_E2=0.008407,
                        gCa_HVAbar_Ca_HVA=0.000994, gNaTs2_tbar_NaTs2_t=0.983955, gIhbar_Ih=8e-5,
                        decay_CaDynamics_E2=210.485284, gCa_LVAstbar_Ca_LVAst=0.000333)

    def fp(name, p, regs, v):
        return FreeParam(name, p, regs, v, v / 10, v * 10)
    free = [
        fp("gNaTs2_t_soma", "gNaTs2_tbar_NaTs2_t", ("soma",), 0.983955),
        fp("gNaTa_t_axon", "gNaTa_tbar_NaTa_t", ("axon",), 3.137968),
        fp("gNap_Et2_axon", "gNap_Et2bar_Nap_Et2", ("axon",), 0.006827),
        fp("gK_Tst_axon", "gK_Tstbar_K_Tst", ("axon",), 0.089259),
        fp("gK_Pst_axon", "gK_Pstbar_K_Pst", ("axon",), 0.973538),
        fp("gSKv3_1_soma", "gSKv3_1bar_SKv3_1", ("soma",), 0.303472),
        fp("gSKv3_1_axon", "gSKv3_1bar_SKv3_1", ("axon",), 1.021945),
        fp("gSK_E2_soma", "gSK_E2bar_SK_E2", ("soma",), 0.008407),
        fp("gSK_E2_axon", "gSK_E2bar_SK_E2", ("axon",), 0.007104),
        fp("gCa_HVA_soma", "gCa_HVAbar_Ca_HVA", ("soma",), 0.000994),
        fp("gCa_LVAst_axon", "gCa_LVAstbar_Ca_LVAst", ("axon",), 0.008752),
        fp("gIh_somabasal", "gIhbar_Ih", ("soma", "basal"), 8e-5),
    ]
    cm = {"soma": 1.0, "axon": 1.0, "basal": 2.0, "apical": 2.0}
    return _assemble(f"ttpc{m.N}", "ttpc", m, placement, base, free, -65.0, 100.0, cm)


# ----------------------------------------------------------------- stimuli
def step_stimulus(amps_nA: Sequence[float], nt: int, dt: float, t_on: float, t_off: float) -> Tuple[np.ndarray, np.ndarray]:
    """Step currents: amp during steps i with t_on <= i*dt < t_off."""
    t = np.arange(nt) * dt
    on = ((t >= t_on - 1e-9) & (t < t_off - 1e-9)).astype(np.float64)
    return np.outer(np.asarray(amps_nA, np.float64), on), np.full(nt, dt)


# ----------------------------------------------------------------- configurations
@dataclasses.dataclass
class Workload:
    name: str
    model: Model
    x: np.ndarray          # free parameters [nsets, nfree]
    amps: np.ndarray       # [nstim, nt]
    dts: np.ndarray        # [nt]

    @property
    def params(self) -> np.ndarray:
        return self.model.allparams(self.x)

    @property
    def nruns(self) -> int:
        return self.x.shape[0] * self.amps.shape[0]

    @property
    def comp_steps(self) -> float:
        return float(self.nruns) * self.model.tree.N * self.dts.size


# Uneven batch: (segments, instances, dt steps) of four cells.
MIXED = ((200, 12000, 10000), (700, 4000, 5000), (1500, 1000, 2000), (3000, 256, 40000))


def mixed_workloads(nsets_scale: float = 1.0, nt_scale: float = 1.0, seed: int = 1003,
                    mechset: str = "mainen") -> List["Workload"]:
    """The uneven batch: pyramidal morphologies of 200, 700, 1,500 and 3,000
    segments with the active cell's channel set (Mainen & Sejnowski: na, kv,
    km, ca, kca, cad, pas, as :func:`active_model`), 12,000 / 4,000 / 1,000 /
    256 instances and 10,000 / 5,000 / 2,000 / 40,000 steps of dt = 0.02 ms,
    one 0.5 nA step from 10 ms to 80 % of each run.  Seeds seed + 10 k per
    morphology.  ``mechset="ttpc"`` builds the same sizes with the TTPC
    mechanisms (:func:`ttpc_model`, dt = 0.025 ms) instead.  The scales
    shrink instances and steps for tests."""
    if mechset not in ("mainen", "ttpc"):
        raise ValueError(f"mixed batches use the mainen or ttpc set, not {mechset!r}")
    out = []
    for k, (n, inst, nt) in enumerate(MIXED):
        if mechset == "mainen":
            model, dt = active_model(seed + 10 * k, n_target=n), 0.02
        else:
            model, dt = ttpc_model(seed + 10 * k, n_target=n), 0.025
        ns = max(1, int(round(inst * nsets_scale)))
        T = max(2, int(round(nt * nt_scale)))
        x = model.sample(ns, seed + 10 * k)
        a, d = step_stimulus([0.5], T, dt, 10.0, 0.8 * T * dt)
        out.append(Workload(f"mixed:{model.name}", model, x, a, d))
    return out


WORKLOADS = ("passive64", "active500", "ttpc1200")
# each workload's step current: (on, off) in ms
STIM_WINDOW = {"passive64": (10.0, 90.0), "active500": (20.0, 180.0), "ttpc1200": (20.0, 280.0)}
_SEEDS = {"passive64": 1000, "active500": 1001, "ttpc1200": 1002}


def workload(name: str, nsets: Optional[int] = None, nt: Optional[int] = None, seed: Optional[int] = None) -> Workload:
    """The benchmark workloads (default seeds 1000, 1001, 1002):
    passive64  passive tree, 64 compartments, 32 instances, 0.5 nA, dt 0.025 ms, 100 ms;
    active500  active ~500-compartment cell, 1024 instances x (0.2, 0.5, 1.0) nA, dt 0.02 ms, 200 ms;
    ttpc1200   TTPC-like ~1200-compartment cell, 8192 instances x 5 amplitudes 0.1-0.9 nA,
               dt 0.025 ms, 300 ms.
    ``nsets`` / ``nt`` shrink a workload for tests."""
    if name not in _SEEDS:
        raise ValueError(f"no workload {name!r}; known: {', '.join(WORKLOADS)}")
    seed = _SEEDS[name] if seed is None else seed
    if name == "passive64":
        model = passive_model(seed)
        n, T, dt, amps = 32, 4000, 0.025, [0.5]
    elif name == "active500":
        model = active_model(seed)
        n, T, dt, amps = 1024, 10000, 0.02, [0.2, 0.5, 1.0]
    else:
        model = ttpc_model(seed)
        n, T, dt, amps = 8192, 12000, 0.025, [0.1, 0.3, 0.5, 0.7, 0.9]
    on, off = STIM_WINDOW[name]
    n = n if nsets is None else nsets
    T = T if nt is None else nt
    x = model.sample(n, seed)
    a, d = step_stimulus(amps, T, dt, on, off)
    return Workload(name, model, x, a, d)
