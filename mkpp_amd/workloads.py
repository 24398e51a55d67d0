"""Seeded synthetic cell states for the benchmark configurations.

Every generator returns a dict with numpy arrays in the C ABI's LayoutLeft
shape: conc (N, cells), rate_parameters (P, cells), temp (cells,),
air_density (cells,), pressure (cells,), plus the method, interval and
tolerances the configuration integrates with.

* chapman_fixed   - chapman, Rodas3, 60 x 60 s fixed steps (seed 1, atol = inf accepts every step)
* small_strato    - Ros4 adaptive, rtol 1e-3, atol 1e3 molec/cm3, 3600 s
* saprc99_urban   - Rodas3 adaptive, rtol 1e-3, atol 1e3 (assumed), 3600 s; the KPP
                    urban base state of the reference's
                    tests/integration/e2e_validation/test_e2e_mechanisms.cpp:100-144 with
                    NOx / VOCs rescaled log-uniformly, photolysis scaled by cos(SZA) and
                    half of the cells at night
* carbon_mixed    - carbon, Rodas3 adaptive, T-dependent rate constants
* saprcnov_urban  - saprcnov with the saprc99 urban setup and matched photolysis rates
* ts1_scenario    - ts1 (mol m-3): the reference runner's binding data, Ros3 at the
                    measurement scenario's accuracy controls, 300 s + 300 s
* gocart_mixed    - gocart: gases, aerosol bins, sulfate ratios across the
                    NH4/NO3/SO4 partition's regimes and cloud water around the
                    gate, Rodas3 adaptive, rtol 1e-3, atol 1e3, 1 h
* batch_mixed     - the batching workload: saprc99 (NOx, VOCs 1e8-1e13, half night),
                    carbon and small_strato (half night) cells, Rodas3 adaptive
                    at rtol 1e-3 for all three in one launch
"""

from __future__ import annotations

import math

import numpy as np

KB = 1.380649e-23  # J/K


def number_density(p_pa, T):
    """Air number density in molecules/cm3."""
    return p_pa / (KB * T) * 1.0e-6


def chapman_fixed(num_cells: int = 4096, seed: int = 1) -> dict:
    rng = np.random.default_rng(seed)
    T = rng.uniform(200.0, 300.0, num_cells)
    p = rng.uniform(1.0e3, 1.0e5, num_cells)
    M = number_density(p, T)
    conc = np.zeros((4, num_cells))
    conc[0] = 10.0 ** rng.uniform(5.0, 8.0, num_cells)  # O
    conc[2] = 10.0 ** rng.uniform(11.0, 13.0, num_cells)  # O3
    conc[3] = M
    conc[1] = 0.2095 * M  # O2 evolves (only AIR/M are fixed)
    jv = rng.uniform(0.0, 1.0e-3, (2, num_cells))
    return {"mechanism": "chapman", "conc": conc, "rate_parameters": jv, "temp": T, "pressure": p,
            "air_density": M, "method": "rodas3", "dt": 60.0, "steps": 60, "seed": 1.0,
            # non-rejecting tolerances: sci = inf makes every error estimate 0,
            # clamped to 1e-10, so each integrate(60 s) is exactly one step
            "atol": math.inf, "rtol": 0.0}


SMALL_STRATO_CENTRES = {"O": 1e7, "O1D": 1e2, "O3": 4e12, "NO": 1e9, "NO2": 1e9}


def small_strato(num_cells: int = 65536, seed: int = 2) -> dict:
    rng = np.random.default_rng(seed)
    names = ["O", "O1D", "O3", "NO", "NO2", "M", "O2"]
    T = rng.uniform(200.0, 260.0, num_cells)
    p = 10.0 ** rng.uniform(2.0, 4.0, num_cells)
    M = number_density(p, T)
    conc = np.zeros((7, num_cells))
    for nm, c in SMALL_STRATO_CENTRES.items():
        conc[names.index(nm)] = c * 10.0 ** rng.uniform(-1.0, 1.0, num_cells)
    conc[names.index("M")] = M
    conc[names.index("O2")] = 0.2095 * M
    sun = rng.uniform(0.0, 1.0, num_cells)
    # host forcing expressions in the mechanism's metadata
    jv = np.stack([2.643e-10 * sun ** 3, 6.120e-04 * sun, 1.070e-03 * sun ** 2, 1.289e-02 * sun])
    return {"mechanism": "small_strato", "conc": conc, "rate_parameters": jv, "temp": T, "pressure": p,
            "air_density": M, "method": "ros4", "dt": 3600.0, "steps": 1, "seed": 1.0e-6,
            "atol": 1.0e3, "rtol": 1.0e-3}


# KPP saprc99 urban initial state in ppm (scaled by CFACTOR = 2.4476e13).
SAPRC99_URBAN_PPM = {
    "NO": 0.1, "NO2": 0.05, "HONO": 0.001, "SO2": 0.05, "HCHO": 0.01121, "CCHO": 0.002316, "RCHO": 0.00172,
    "ACET": 0.00507, "MEK": 0.00326, "HCOOH": 0.000677, "MEOH": 0.00589, "CCO_OH": 0.00116, "RCO_OH": 0.000392,
    "GLY": 0.000121, "MGLY": 8.37e-05, "CRES": 0.00056, "BALD": 7.51e-05, "ISOPROD": 8.93e-05,
    "METHACRO": 0.0013, "PROD2": 0.00193, "ETHENE": 0.0189, "ISOPRENE": 0.000433, "ALK1": 0.01167,
    "ALK2": 0.0188, "ALK3": 0.0469, "ALK4": 0.0417, "ALK5": 0.0306, "ARO1": 0.0118, "ARO2": 0.00874,
    "OLE1": 0.0104, "OLE2": 0.00797, "TERP": 0.00082, "PHEN": 0.000606, "XC": 0.2, "O3P": 7.843e-09,
}
SAPRC99_VOCS = ["ALK1", "ALK2", "ALK3", "ALK4", "ALK5", "ARO1", "ARO2", "OLE1", "OLE2", "ETHENE", "ISOPRENE",
                "TERP", "HCHO", "CCHO", "RCHO", "ACET", "MEK"]
# Photolysis J-set at SUN = 1 (test_e2e_mechanisms.cpp:48-86): A * factor / 60.
SAPRC99_J1 = [6.69e-1, 1.59e0, 1.50e+1, 3.76e-2, 4.19e-3, 1.27e-1, 1.60e-2, 5.40e-5, 4.69e-4, 5.64e-4, 2.32e-3,
              3.15e-3, 4.16e-4, 1.40e-3, 4.16e-5, 9.49e-4 * 1.50e-1, 3.94e-4, 3.94e-4, 8.93e-3, 1.81e-1 * 6.00e-3,
              1.10e-2, 1.90e-2, 6.22e-2 * 5.00e-2, 3.32e-2 * 4.10e-3, 3.32e-2 * 2.10e-3, 3.32e-2 * 4.10e-3,
              9.49e-4 * 2.00e-2, 2.35e-4, 2.06e-1 * 3.65e-1, 3.32e-2 * 7.28e0]


def saprc99_urban(num_cells: int = 262144, seed: int = 3, species=None, nox_range=(1e9, 1e12),
                  voc_range=(1e9, 1e12), night_fraction=0.5, jset=None, mechanism="saprc99") -> dict:
    if species is None:
        from .gen.model import load_mechanism

        species = load_mechanism("saprc99").names
    idx = {s: i for i, s in enumerate(species)}
    rng = np.random.default_rng(seed)
    n = len(species)
    cf = 2.4476e13
    T = rng.uniform(270.0, 310.0, num_cells)
    p = np.full(num_cells, 101325.0)
    air = number_density(p, T)
    conc = np.zeros((n, num_cells))
    for nm, ppm in SAPRC99_URBAN_PPM.items():
        if nm in idx:  # a species the mechanism lacks is left out
            conc[idx[nm]] = ppm * cf
    nox = 10.0 ** rng.uniform(math.log10(nox_range[0]), math.log10(nox_range[1]), num_cells)
    conc[idx["NO"]] = nox * 2.0 / 3.0
    conc[idx["NO2"]] = nox / 3.0
    for v in SAPRC99_VOCS:
        draw = 10.0 ** rng.uniform(math.log10(voc_range[0]), math.log10(voc_range[1]), num_cells)
        if v in idx:
            conc[idx[v]] = draw
    conc[idx["AIR"]] = air
    conc[idx["M"]] = air
    conc[idx["O2"]] = 0.209 * air
    conc[idx["N2"]] = 0.78 * air
    conc[idx["H2O"]] = 0.01 * air
    conc[idx["CH4"]] = 1.8e-6 * air
    if "H2" in idx:
        conc[idx["H2"]] = 5.0e-7 * air
    sza = rng.uniform(0.0, 90.0, num_cells)
    sun = np.cos(np.radians(sza))
    night = rng.uniform(0.0, 1.0, num_cells) < night_fraction
    sun[night] = 0.0
    jv = np.outer(np.array(SAPRC99_J1 if jset is None else jset) / 60.0, sun)
    return {"mechanism": mechanism, "conc": conc, "rate_parameters": jv, "temp": T, "pressure": p,
            "air_density": air, "method": "rodas3", "dt": 3600.0, "steps": 1, "seed": 1.0e-6,
            "atol": 1.0e3, "rtol": 1.0e-3, "night": night}


def carbon_mixed(num_cells: int = 4096, seed: int = 4) -> dict:
    rng = np.random.default_rng(seed)
    names = ["CH4", "CO", "PCOfromCH4", "PCOfromNMVOC", "LCH4byOH", "LCH4byCl", "LCObyOH", "FixedOH", "FixedCl",
             "DummyCH4", "DummyNMVOC"]
    T = rng.uniform(220.0, 310.0, num_cells)
    p = 10.0 ** rng.uniform(3.0, 5.0, num_cells)
    air = number_density(p, T)
    conc = np.zeros((11, num_cells))
    conc[names.index("CH4")] = 1.8e-6 * air
    conc[names.index("CO")] = 10.0 ** rng.uniform(11.0, 13.0, num_cells)
    conc[names.index("FixedOH")] = 10.0 ** rng.uniform(5.0, 7.0, num_cells)
    conc[names.index("FixedCl")] = 10.0 ** rng.uniform(2.0, 4.0, num_cells)
    conc[names.index("DummyCH4")] = 1.0
    conc[names.index("DummyNMVOC")] = 1.0
    return {"mechanism": "carbon", "conc": conc, "rate_parameters": np.zeros((0, num_cells)), "temp": T,
            "pressure": p, "air_density": air, "method": "rodas3", "dt": 3600.0, "steps": 1, "seed": 1.0e-6,
            "atol": 1.0e3, "rtol": 1.0e-3}


def batch_mixed(num_saprc99: int = 131072, num_carbon: int = 65536, num_small_strato: int = 65536,
                seed: int = 4) -> list:
    """Three workloads of deliberately uneven stiffness for one persistent launch.

    saprc99 urban/rural cells with NOx and VOCs log-uniform over [1e8, 1e13]
    and half of them at night; carbon cells (no photolysis); small_strato cells
    with half of them at night.  All integrate 3600 s with Rodas3 at rtol 1e-3,
    atol 1e3.  Accepted steps per cell then range from 9 (the floor set by the
    controller: growth at most x6 per step from the 1e-6 * 3600 s seed) to
    about 70 (sunlit saprc99 cells)."""
    sap = saprc99_urban(num_saprc99, seed=seed, nox_range=(1e8, 1e13), voc_range=(1e8, 1e13))
    car = carbon_mixed(num_carbon, seed=seed + 1)
    sst = small_strato(num_small_strato, seed=seed + 2)
    rng = np.random.default_rng(seed + 3)
    night = rng.uniform(0.0, 1.0, num_small_strato) < 0.5
    sst["rate_parameters"][:, night] = 0.0
    sst["night"] = night
    out = []
    for w in (sap, car, sst):
        out.append(dict(w, method="rodas3", dt=3600.0, steps=1, seed=1.0e-6, atol=1.0e3, rtol=1.0e-3))
    return out


def saprcnov_jset() -> list:
    """Photolysis rates (min^-1, overhead sun) for saprcnov's slots: the
    saprc99 set's rate of the same photolysed species and occurrence (the
    n-th photolysis of NO2 takes saprc99's n-th NO2 rate), else the median of
    the saprc99 set."""
    from .gen.model import load_mechanism

    def keyed(m):
        seen, out = {}, []
        for r in m.reactions:
            if r.kind.upper().startswith("PHOTO"):
                sp_ = m.names[r.reactants[0][0]]
                seen[sp_] = seen.get(sp_, 0) + 1
                out.append((sp_, seen[sp_]))
        return out

    ref = dict(zip(keyed(load_mechanism("saprc99")), SAPRC99_J1))
    median = float(np.median(SAPRC99_J1))
    return [ref.get(key, median) for key in keyed(load_mechanism("saprcnov"))]


def saprcnov_urban(num_cells: int = 65536, seed: int = 6, night_fraction=0.5) -> dict:
    """saprcnov with the saprc99 urban setup (species saprcnov lacks are left
    out, its extra species start at 0) and saprcnov_jset's photolysis rates;
    Rodas3 adaptive, rtol 1e-3, atol 1e3, 3600 s."""
    from .gen.model import load_mechanism

    return saprc99_urban(num_cells, seed=seed, species=load_mechanism("saprcnov").names,
                         night_fraction=night_fraction, jset=saprcnov_jset(), mechanism="saprcnov")


def saprc99_diff(num_cells: int = 65536, seed: int = 5) -> dict:
    """The differentiation workload: saprc99 urban/rural cells (as
    saprc99_urban, seed 5) plus ``target_initial``, the same states with every
    variable species scaled by (1 + 0.1 N(0, 1)); integrating it gives the
    target of the loss sum over cells of (O3 - O3*)^2 + (NO2 - NO2*)^2 after 1 h."""
    from .gen.model import load_mechanism

    mech = load_mechanism("saprc99")
    w = saprc99_urban(num_cells, seed=seed, species=mech.names)
    rng = np.random.default_rng(seed + 100)
    target = w["conc"] * (1.0 + 0.1 * rng.standard_normal(w["conc"].shape))
    fixed = list(mech.fixed)
    target[fixed] = w["conc"][fixed]
    w["target_initial"] = np.maximum(target, 0.0)
    w["loss_species"] = [mech.names.index("O3"), mech.names.index("NO2")]
    return w


def ts1_scenario(num_cells: int = 1024, seed: int | None = 7, spread: float = 0.1) -> dict:
    """The ts1 measurement scenario (mol m-3; 210 species, 547 rate-parameter
    slots): the initial state, rate parameters, temperature and air density the
    reference's solver-comparison runner binds (runners/mkpp/main.cpp:206-215,
    326-342; data in mechanisms/ts1/scenario.json), Ros3, the scenario's
    accuracy controls (rtol 1e-6, atol per species).  ``rate_parameters`` is
    the first segment's; ``segments`` lists (seconds, photolysis scale of slots
    0..122) for the runner's 300 s + 300 s schedule.

    The runner gives every cell the same state.  With ``seed`` set, each cell's
    initial state and photolysis rates are scaled by exp(spread * N(0, 1)) so
    that the cells differ; seed None keeps them identical."""
    import json
    import os

    d = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "mechanisms", "ts1", "scenario.json")))
    y0 = np.asarray(d["initial_state"], dtype=np.float64)
    rp0 = np.asarray(d["reaction_parameters"], dtype=np.float64)
    conc = np.repeat(y0[:, None], num_cells, axis=1)
    rp = np.repeat(rp0[:, None], num_cells, axis=1)
    photo = np.asarray(d["photolysis_slots"])
    if seed is not None:
        rng = np.random.default_rng(seed)
        fixed = d["species"].index("M")
        scale = np.exp(spread * rng.standard_normal(conc.shape))
        scale[fixed] = 1.0
        conc = conc * scale
        rp[photo] = rp[photo] * np.exp(spread * rng.standard_normal((len(photo), num_cells)))
    T = np.full(num_cells, float(d["temperature"]))
    rho = np.full(num_cells, float(d["air_density"]))
    return {"mechanism": "ts1", "conc": np.ascontiguousarray(conc), "rate_parameters": np.ascontiguousarray(rp),
            "temp": T, "pressure": np.full(num_cells, 101319.9), "air_density": rho, "method": "ros3",
            "dt": 300.0, "steps": 1, "seed": 1.0e-6, "atol": np.asarray(d["atol"]), "rtol": float(d["rtol"]),
            "photolysis_slots": photo,
            "segments": [(s["seconds"], s["photolysis_scale"]) for s in d["segments"]]}


GOCART_CENTRES = {
    # gases, molecules cm-3
    "OH": 1e6, "NO3": 1e7, "H2O2": 1e10, "DMS": 1e9, "SO2": 1e10, "SO4": 1e10, "HNO3": 1e10, "NH3": 1e10,
    "NH4a": 1e10, "NO3an1": 3e9, "NO3an2": 3e9, "NO3an3": 3e9,
    # carbonaceous aerosol tracers
    "BC1": 1e9, "BC2": 1e9, "OC1": 1e9, "OC2": 1e9,
}


def gocart_mixed(num_cells: int = 65536, seed: int = 9) -> dict:
    """gocart cells: gases log-uniform within a decade of GOCART_CENTRES,
    dust and sea-salt bins (the heterogeneous reactions' second reactant, read
    as uptake surfaces with v_gas = S_a = 1) log-uniform in 1e-7..1e-5, the
    sulfate ratio (NH3 + NH4a) / (SO2 + SO4) log-uniform in 0.2..20 (both
    sides of the partition's regime boundaries at 1 and 2), T in 250..310 K,
    H2O2 photolysis up to 1e-5 s-1 (the mechanism's host forcing), and cloud
    liquid water 0 in a third of the cells and log-uniform in 1e-7..1e-5
    (around the 1e-6 gate of SO2 + H2O2) elsewhere.  Rodas3 adaptive, rtol
    1e-3, atol 1e3, 1 h."""
    from .gen.model import load_mechanism

    mech = load_mechanism("gocart")
    names = mech.names
    rng = np.random.default_rng(seed)
    conc = np.zeros((len(names), num_cells))
    for i, nm in enumerate(names):
        if nm in GOCART_CENTRES:
            conc[i] = GOCART_CENTRES[nm] * 10.0 ** rng.uniform(-1.0, 1.0, num_cells)
        else:
            conc[i] = 10.0 ** rng.uniform(-7.0, -5.0, num_cells)
    cs = conc[names.index("SO2")] + conc[names.index("SO4")]
    cn = cs * 10.0 ** rng.uniform(np.log10(0.2), np.log10(20.0), num_cells)
    f = rng.uniform(0.1, 0.9, num_cells)
    conc[names.index("NH3")], conc[names.index("NH4a")] = f * cn, (1.0 - f) * cn
    T = rng.uniform(250.0, 310.0, num_cells)
    p = rng.uniform(5e4, 1e5, num_cells)
    clw = np.where(rng.uniform(0.0, 1.0, num_cells) < 1.0 / 3.0, 0.0, 10.0 ** rng.uniform(-7.0, -5.0, num_cells))
    jv = rng.uniform(0.0, 1e-5, (1, num_cells))
    return {"mechanism": "gocart", "conc": conc, "rate_parameters": jv, "temp": T, "pressure": p,
            "air_density": number_density(p, T), "cloud_water": clw, "method": "rodas3", "dt": 3600.0,
            "steps": 1, "seed": 1.0e-6, "atol": 1.0e3, "rtol": 1.0e-3}
