"""The GCMC move-energy path on its own workload: CO2 in the hexagonal 13/13/15 Å cell (gamma
120) replicated 3x3x2, T = 298.15 K, fugacities 1e4, 1e5 and 1e6 Pa in one batch, moves
insertion / deletion / translation (Gaussian, s.d. 1 Å) / rotation (step 0.5) at 1/4 each,
guest-guest LJ + Ewald with incremental structure factors (src/gcmc.jl:161-337).

The 300-atom cell the workload names is non-porous (no insertion is ever accepted), so the batch
also holds a 40-atom cell of the same shape in which every chain accepts tens of insertions.
Every chain is replayed on the FP64 oracle from the same streams: occupancy sequences equal,
per-cycle energies within 1e-8, and the GCMCResult fields (loading, energy and their block
errors, fluctuation q_st, its jackknife error, corr(U, N)) against the oracle's block_mean_sem
and fluctuation_qst of the oracle's own samples."""
from __future__ import annotations

from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import oracle as O
import pureadsorb_amd as pa
from pureadsorb_amd import synth
from conftest import oracle_for

pytestmark = pytest.mark.gpu

T = 298.15
FUG = [1.0e4, 1.0e5, 1.0e6]
NCYC, NAUDIT, NBLOCKS, SEED = 600, 200, 10, 7


def test_config4_gcmc_matches_the_oracle_replay():
    ff = synth.synthetic_forcefield(12.0)
    g = synth.co2_guest(ff)
    ew = pa.EwaldParams(12.0, 1e-6)
    lit = pa.replicate(synth.synthetic_framework(13.0, 13.0, 15.0, 90, 90, 120, 300, 42, charge_sd=0.5), (3, 3, 2))
    por = pa.replicate(synth.synthetic_framework(13.0, 13.0, 15.0, 90, 90, 120, 40, 42, charge_sd=0.5), (3, 3, 2))
    fws = [lit] * 3 + [por] * 3
    fug = FUG * 2
    b = pa.FrameworkBatch(fws, ff, g, ew, fullk=True)
    st = pa.SystemState(b, g, [0] * 6, T=T, seed=SEED, capacities=[400] * 6)
    print(f"config 4: 6 chains (3 on the 300-atom cell, 3 on the 40-atom cell) x {NCYC} cycles", flush=True)
    res = pa.run_gcmc(st, T=T, n_warmup=0, n_production=NCYC, n_audit=NAUDIT, step_trans=1.0, step_rot=0.5,
                      fugacity=fug, p_translation=0.25, p_rotation=0.25, p_exchange=0.5, min_cycle_length=1,
                      nblocks=NBLOCKS)
    osys = {id(lit): oracle_for(lit, ff, g, ew), id(por): oracle_for(por, ff, g, ew)}

    def replay(n):
        m = O.OracleMC(osys[id(fws[n])], 400, O.splitmix64(SEED, n + 1))
        e, occ = [], []
        for c in range(NCYC):
            assert m.cycle((0.25, 0.25, 0.0, 0.5), 1, 1.0, 0.5, fug[n], T) > 0
            if (c + 1) % NAUDIT == 0:
                m.audit()
            s = m.scalars()
            e.append(s["energy"])
            occ.append(s["occupancy"])
        acc, att = m.counts()
        return np.array(e), np.array(occ, dtype=np.float64), acc, att

    with ThreadPoolExecutor(6) as ex:
        reps = list(ex.map(replay, range(6)))
    kT = pa.KB * T
    for n, (r, (e_or, n_or, acc, att)) in enumerate(zip(res, reps)):
        e_dev, n_dev = pa.gcmc_samples(st, n, NCYC)
        print(f"chain {n} ({'300' if n < 3 else '40'} atoms, {fug[n]:.0e} Pa): loading {r.loading:.3f} +- "
              f"{r.loading_err:.3f}, q_st {r.q_st:.4f} +- {r.q_st_err:.4f} eV, accepted {list(r.accepted)}, "
              f"oracle accepted {acc.tolist()}", flush=True)
        assert np.array_equal(n_dev, n_or)
        assert np.allclose(e_dev, e_or, rtol=1e-8, atol=1e-8)
        assert tuple(int(x) for x in acc) == r.accepted and tuple(int(x) for x in att) == r.attempted
        lm, le = O.block_mean_sem(n_or, NBLOCKS)
        em, ee = O.block_mean_sem(e_or, NBLOCKS)
        assert r.loading == pytest.approx(lm, abs=1e-12) and r.loading_err == pytest.approx(le, abs=1e-12)
        assert r.energy == pytest.approx(em, rel=1e-8, abs=1e-8)
        assert r.energy_err == pytest.approx(ee, rel=1e-6, abs=1e-9)
        assert r.max_occupancy == int(n_or.max()) and r.ncycles == NCYC
        if n < 3:
            # the non-porous cell: nothing is ever inserted, so var(N) = 0 and q_st is undefined
            assert r.accepted[3] == 0 and r.loading == 0.0 and np.isnan(r.q_st)
            continue
        assert r.accepted[3] >= 20, r.accepted
        with np.errstate(all="ignore"):
            q, qe, c = O.fluctuation_qst(e_or, n_or, kT, NBLOCKS)
        assert r.q_st == pytest.approx(q, rel=1e-6, abs=1e-9)
        assert r.q_st_err == pytest.approx(qe, rel=1e-4, abs=1e-9)
        assert r.corr_UN == pytest.approx(c, rel=1e-6, abs=1e-9)
