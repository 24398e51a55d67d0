"""The wave-engine tables (row-displacement LU storage, right-looking pivot
schedule, flux/partial/W-entry term lists) reproduce the oracle's F, W,
factorization and solve when interpreted on the CPU exactly as the device
kernel interprets them, and every index stays inside its buffer."""

import numpy as np
import pytest

from mkpp_amd.gen import emit_wave, plan as planmod
from mkpp_amd.gen.model import load_mechanism
from mkpp_amd.gen.plan import LAW_PARAMS
from oracle.oracle import Oracle


@pytest.fixture(scope="module", params=["saprc99", "saprcnov", "saprc99_mini"])
def case(request):
    mech = load_mechanism(request.param)
    p = planmod.build(mech)
    t = emit_wave.build_tables(mech, p)
    return request.param, mech, p, t


def test_indices_in_bounds(case):
    _, _, _, t = case
    nv, ny, span, ng, npart = t["nv"], t["ny"], t["span"], t["ng"], t["npart"]
    # R may be negative; every address a row uses (R[i] + its columns) must not be
    for k in range(nv):
        assert t["R"][k] + k >= 0
    assert min(t["waddr"]) >= 0
    assert max(t["waddr"]) < span and max(t["wpi"] or [0]) < npart
    for k in range(nv):  # pivot and U(k) addresses of every row used by the schedule
        assert t["R"][k] + k < span
        for r in t["lrow"][t["lofs"][k]:t["lofs"][k + 1]]:
            assert 0 <= t["R"][r] + k < span
            for j in t["ucol"][t["uofs"][k]:t["uofs"][k + 1]]:
                assert 0 <= t["R"][r] + j < span and 0 <= t["R"][k] + j < span
    assert max(max(f) for f in t["gfac"]) < ny and max(max(s) for s in t["gpow_s"]) < ny
    assert max(t["fg"]) < ng and max(t["fc"]) < len(t["coefs"]) and max(t["wci"] or [0]) < len(t["coefs"])
    assert max(t["lcnt"]) <= 64 and t["lofs"][-1] == len(t["lrow"]) and t["uofs"][-1] == len(t["ucol"])
    assert t["poff"][-1] == npart and len(t["coefs"]) <= 255
    # row displacement: no two structural non-zeros share an address
    assert len(t["waddr"]) == len(set(t["waddr"]))
    assert ng <= 255 and nv <= 128  # uint8 group / row indices
    # row -> (bank, lane) slot map is a bijection onto the used slots
    assert len(set(t["slot"])) == nv and all(t["rowof"][t["slot"][r]] == r for r in range(nv))
    assert sum(1 for r in t["rowof"] if r != 255) == nv and max(t["slot"]) < 64 * t["nb"]
    # adjoint tables: group -> variable rows, species slot -> partial slots
    assert t["gnvofs"][-1] == len(t["gnvr"]) and max(t["gnvr"]) < nv and max(t["gnvc"]) < len(t["coefs"])
    assert len(t["spofs"]) == nv + t["nf"] + 1 and t["spofs"][-1] == len(t["spi"]) == npart
    assert sorted(t["spi"]) == list(range(npart))  # every partial slot scatters to exactly one species
    assert -32768 <= min(t["R"]) and max(t["R"]) < 32768 and span < 65536  # int16 R, uint16 addresses


def _rate(law, P, T, M, jv):
    if law == 0:
        return jv
    if law == 1:
        k = P[0] * ((T / 300.0) ** P[1] if P[1] != 0.0 else 1.0)
        return k * np.exp(P[2] / T) if P[2] != 0.0 else k
    if law == 2:
        k0 = P[0] * np.exp(P[2] / T) * (T / 300.0) ** P[1]
        ki = P[3] * np.exp(P[5] / T) * (T / 300.0) ** P[4]
        K0 = k0 * M
        r = K0 / ki
        return K0 / (1 + r) * P[6] ** (1.0 / (1.0 + np.log10(r) ** 2))
    if law == 3:
        K0 = P[0] * np.exp(P[1] / T)
        K2 = P[2] * np.exp(P[3] / T)
        K3 = P[4] * np.exp(P[5] / T) * M
        return K0 + K3 / (1 + K3 / K2)
    if law == 4:
        return P[0] * np.exp(P[1] / T) + P[2] * np.exp(P[3] / T) * M
    if law == 5:
        return 0.25 * P[0]
    return P[0]


def test_emulated_wave_engine_matches_oracle(case):
    name, mech, p, t = case
    o = Oracle(name)
    rng = np.random.default_rng(3)
    n, nv = t["n"], t["nv"]
    y = 10.0 ** rng.uniform(7, 12, n)
    tb = t["third_body"]
    M = y[tb] if tb >= 0 else 2.4e19
    T = 287.0
    jv = rng.uniform(0, 1e-3, max(t["nph"], 1))
    k = np.array([_rate(t["law"][g], t["params"][g], T, M, jv[t["photo"][g]] if t["law"][g] == 0 else 0.0)
                  for g in range(t["ng"])])
    kr, _ = o.rate_constants(y, jv, T)
    np.testing.assert_allclose(k, kr, rtol=1e-13)
    # ylds: variable (logical), fixed, ONE
    ylds = np.zeros(t["ny"])
    for v, ph in enumerate(t["var_phys"]):
        ylds[v] = 1.0 if p["unit_value"][ph] else y[ph]
    for q, ph in enumerate(t["fixed_phys"]):
        ylds[nv + q] = y[ph]
    ylds[t["one"]] = 1.0
    coef = np.array(t["coefs"])
    # F
    flux = np.array([k[g] * (((1.0 * ylds[f[0]]) * ylds[f[1]]) * ylds[f[2]]) for g, f in enumerate(t["gfac"])])
    F = np.zeros(nv)
    for v in range(nv):
        acc = 0.0
        for tt in range(t["fofs"][v], t["fofs"][v + 1]):
            acc = acc + coef[t["fc"][tt]] * flux[t["fg"][tt]]
        F[v] = acc
    Fr = o.rates(y, jv, T)
    np.testing.assert_allclose(F, Fr[t["var_phys"]], rtol=1e-12, atol=1e-300)
    # W into row-displacement storage
    part = np.zeros(t["npart"])
    for g in range(t["ng"]):
        np_ = t["gnpow"][g]
        s, q = t["gpow_s"][g], t["gpow_p"][g]
        for e in range(np_):
            v = k[g] * q[e] * (ylds[s[e]] ** (q[e] - 1) if q[e] > 1 else 1.0)
            for o_ in range(np_):
                if o_ != e:
                    v = v * ylds[s[o_]] ** q[o_]
            part[t["poff"][g] + e] = v
    inv_g_dt = 1.0e5  # small step: W well conditioned, so solve orders agree to rounding
    lu = np.full(t["span"], np.nan)
    for e, addr in enumerate(t["waddr"]):
        acc = 0.0
        for tt in range(t["wofs"][e], t["wofs"][e + 1]):
            acc = acc + coef[t["wci"][tt]] * part[t["wpi"][tt]]
        lu[addr] = (inv_g_dt if t["wdiag"][e] else 0.0) - acc
    # oracle W in its slot order -> compare through the var LU locations
    Js = o.jacobian_slots(y, jv, T)
    w = np.zeros(len(p["w_locations"]))
    for s_, (i, j) in enumerate(p["w_locations"]):
        base = inv_g_dt if i == j else 0.0
        js = p["w_from_j"][s_]
        w[s_] = base - Js[js] if js >= 0 else base
    inv = {ph: lg for lg, ph in enumerate(p["permutation"])}
    vlog = {lg: v for v, lg in enumerate([lg for lg in range(n) if p["permutation"][lg] in t["var_phys"]])}
    R = t["R"]
    for s_, (i, j) in enumerate(p["w_locations"]):
        if i in vlog and j in vlog:
            assert abs(lu[R[vlog[i]] + vlog[j]] - w[s_]) <= 1e-12 * abs(w[s_]) + 1e-300
    # right-looking factorization, then compare L and U with the oracle's plan
    for kk in range(nv):
        d = lu[R[kk] + kk]
        for r in t["lrow"][t["lofs"][kk]:t["lofs"][kk + 1]]:
            l = lu[R[r] + kk] / d
            lu[R[r] + kk] = l
            for j in t["ucol"][t["uofs"][kk]:t["uofs"][kk + 1]]:
                lu[R[r] + j] = lu[R[r] + j] - l * lu[R[kk] + j]
    LUo = o.factorize(w)
    for s_, (i, j) in enumerate(p["lu_locations"]):
        if i in vlog and j in vlog:
            a, b = lu[R[vlog[i]] + vlog[j]], LUo[s_]
            assert abs(a - b) <= 1e-10 * abs(b) + 1e-300, (i, j, a, b)
    # column-oriented solve with the masks
    rhs = rng.uniform(-1, 1, nv)
    wv = rhs.copy()
    for kk in range(nv):
        for r in range(nv):
            if (t["fmask"][kk][r // 64] >> (r % 64)) & 1:
                wv[r] -= lu[R[r] + kk] * wv[kk]
    x = np.zeros(nv)
    for kk in reversed(range(nv)):
        x[kk] = wv[kk] / lu[R[kk] + kk]
        for r in range(nv):
            if (t["bmask"][kk][r // 64] >> (r % 64)) & 1:
                wv[r] -= lu[R[r] + kk] * x[kk]
    # transposed solve W^T x = b as the generated tforward / tbackward run it:
    # lane j of pivot k reads R[k] + j (row k of U, then row k of L)
    bt = rng.uniform(-1, 1, nv)
    zb = bt.copy()
    z = np.zeros(nv)
    for kk in range(nv):
        z[kk] = zb[kk] / lu[R[kk] + kk]
        for j in t["ucol"][t["uofs"][kk]:t["uofs"][kk + 1]]:
            zb[j] -= lu[R[kk] + j] * z[kk]
    for kk in reversed(range(nv)):
        for j in range(kk):
            if kk in t["lrow"][t["lofs"][j]:t["lofs"][j + 1]]:
                z[j] -= lu[R[kk] + j] * z[kk]
    Wd = np.zeros((nv, nv))
    for s_, (i, j) in enumerate(p["w_locations"]):
        if i in vlog and j in vlog:
            Wd[vlog[i], vlog[j]] = w[s_]
    zr = np.linalg.solve(Wd.T, bt)
    np.testing.assert_allclose(z, zr, rtol=1e-9, atol=1e-8 * np.abs(zr).max())
    full_rhs = np.zeros(n)
    lg_of_v = [lg for lg in range(n) if p["permutation"][lg] in t["var_phys"]]
    for v, lg in enumerate(lg_of_v):
        full_rhs[lg] = rhs[v]
    xo = o.solve(LUo, full_rhs)
    np.testing.assert_allclose(x, xo[lg_of_v], rtol=1e-9, atol=1e-300)


def test_chunked_f_tables_match_oracle(case):
    """CF_TERM / CF_PART decoded as wave_eval_f decodes them: chunk partials
    into scratch slots, then each row's partials in chunk order."""
    import copy

    name, mech, p, t = case
    t = copy.deepcopy(t)
    ct = emit_wave.build_code_tables(t)  # may append 0.0 to t["coefs"]
    o = Oracle(name)
    rng = np.random.default_rng(11)
    n, nv, ng, nb = t["n"], t["nv"], t["ng"], t["nb"]
    tf, nfp, pbase, maxcp = ct["tf"], ct["nfp"], ct["pbase"], ct["fmaxcp"]
    nscr = t["nscr"]
    assert nscr >= max(ng, t["npart"]) + 1
    # bounds: every written partial slot inside the scratch and row-sum areas,
    # every term inside the flux slots and the coefficient table
    assert pbase == ng + 1 and pbase + 64 * nfp <= nscr + nv + 1
    assert len(ct["fterm"]) == nfp * 64 * (tf // 2) and len(ct["fpart"]) == nb * 64 * maxcp
    terms = [(w >> s) & 0xFFFF for w in ct["fterm"] for s in (0, 16)]
    assert max(x & 0xFF for x in terms) <= ng and max(x >> 8 for x in terms) < len(t["coefs"])
    assert all(ng <= i < pbase + 64 * nfp for i in ct["fpart"])
    assert all(1 <= m <= maxcp for m in ct["fmaxc"])
    y = 10.0 ** rng.uniform(7, 12, n)
    tb = t["third_body"]
    M = y[tb] if tb >= 0 else 2.4e19
    T = 291.0
    jv = rng.uniform(0, 1e-3, max(t["nph"], 1))
    k = np.array([_rate(t["law"][g], t["params"][g], T, M, jv[t["photo"][g]] if t["law"][g] == 0 else 0.0)
                  for g in range(ng)])
    ylds = np.zeros(t["ny"])
    for v, ph in enumerate(t["var_phys"]):
        ylds[v] = 1.0 if p["unit_value"][ph] else y[ph]
    for q, ph in enumerate(t["fixed_phys"]):
        ylds[nv + q] = y[ph]
    ylds[t["one"]] = 1.0
    coef = np.array(t["coefs"])
    scr = np.full(nscr + nv + 1, np.nan)
    for g, f in enumerate(t["gfac"]):
        scr[g] = k[g] * (((1.0 * ylds[f[0]]) * ylds[f[1]]) * ylds[f[2]])
    scr[ng] = 0.0
    for pp in range(nfp):
        for lane in range(64):
            acc = 0.0
            for q in range(tf):
                x = terms[(pp * 64 + lane) * tf + q]
                acc = acc + coef[x >> 8] * scr[x & 0xFF]
            scr[pbase + pp * 64 + lane] = acc
    F = np.zeros(nv)
    for v in range(nv):
        b, lane = divmod(t["slot"][v], 64)
        lst = ct["fpart"][(b * 64 + lane) * maxcp:(b * 64 + lane + 1) * maxcp]
        s = scr[lst[0]]
        for c in range(1, ct["fmaxc"][b]):
            s = s + scr[lst[c]]
        F[v] = s
        assert all(i == ng for i in lst[ct["fmaxc"][b]:])  # entries past FMAXC are padding only
    Fr = o.rates(y, jv, T)
    np.testing.assert_allclose(F, Fr[t["var_phys"]], rtol=1e-12, atol=1e-300)


def test_chunked_scatter_tables(case):
    """CS_TERM / CS_PART / CS_FIX decoded as the adjoint's scatter decodes
    them give each species slot the sum of its partial slots (every slot used
    once), and every index stays inside the areas the kernel gives them."""
    import copy

    _, _, _, t = case
    t = copy.deepcopy(t)
    ct = emit_wave.build_code_tables(t)
    nv, nf, nb, npart, ng, ny = t["nv"], t["nf"], t["nb"], t["npart"], t["ng"], t["ny"]
    ts, nsp, szero = ct["ts"], ct["nsp"], ct["szero"]
    assert szero == 64 * nsp
    terms = [(w >> s) & 0xFFFF for w in ct["sterm"] for s in (0, 16)]
    assert len(terms) == nsp * 64 * ts and max(terms) <= npart
    assert all(0 <= i <= szero for i in ct["spart"] + ct["sfix"])
    rng = np.random.default_rng(5)
    scr = np.append(rng.standard_normal(npart), 0.0)
    cp = np.full(szero + 1, np.nan)
    for j in range(nsp * 64):
        acc = 0.0
        for q in range(ts):
            acc += scr[terms[j * ts + q]]
        cp[j] = acc
    cp[szero] = 0.0
    want = [sum(scr[i] for i in t["spi"][t["spofs"][l]:t["spofs"][l + 1]]) for l in range(nv + nf)]
    for v in range(nv):
        b, lane = divmod(t["slot"][v], 64)
        lst = ct["spart"][(b * 64 + lane) * ct["smaxcp"]:(b * 64 + lane + 1) * ct["smaxcp"]]
        assert all(i == szero for i in lst[ct["smaxc"][b]:])
        got = cp[lst[0]]
        for c in range(1, ct["smaxc"][b]):
            got += cp[lst[c]]
        assert abs(got - want[v]) <= 1e-12 * (1 + abs(want[v]))
    for f in range(nf):
        lst = ct["sfix"][f * ct["sfmaxcp"]:(f + 1) * ct["sfmaxcp"]]
        got = sum(cp[lst[c]] for c in range(ct["sfmaxc"]))
        assert abs(got - want[nv + f]) <= 1e-12 * (1 + abs(want[nv + f]))
    # every partial slot reaches exactly one species once
    used = sorted(i for i in terms if i != npart)
    assert used == list(range(npart))


def test_law_parameter_order():
    assert LAW_PARAMS["TROE"] == ["A0", "B0", "C0", "A1", "B1", "C1", "Fc"]
    assert LAW_PARAMS["EP2"] == ["A0", "C0", "A2", "C2", "A3", "C3"]
    assert LAW_PARAMS["EP3"] == ["A1", "C1", "A2", "C2"]
