"""Serving loop: HBM-budget admission and continuous batching of MD
segments; every job's result is the same alone and in the mix (bit for
bit: per-system argument structs and fixed-order per-system sums make a
system's arithmetic independent of its batch companions)."""
import numpy as np
import pytest

from mdynamix_amd import systems
from mdynamix_amd.engine import ForceEngine, MdxError
from mdynamix_amd.serve import MDJob, Server

pytestmark = pytest.mark.gpu


def _jobs(nsteps):
    spec = [(1500, 10, 0.3, 11), (2600, 24, 0.8, 12), (1800, 16, 0.05, 13), (3400, 40, 1.0, 14)]
    out = []
    for n, ns, lam, seed in spec:
        s = systems.ee_job(n, ns, lam, seed, rigid_solute=True)
        out.append(MDJob(s, nsteps, systems.maxwell_velocities(s, seed)))
    return out


def _alone(job):
    eng = ForceEngine(job.topo, job.prm)
    eng.set_positions(job.system.pos)
    eng.md_init(job.md, np.stack([v for v in job.vel], 1))
    tr = eng.md_run(job.nsteps)
    pos, vel = eng.md_state()
    eng.close()
    return pos, vel, int(round(tr[:, 4].sum())), tr[-1]


def test_served_jobs_match_alone():
    jobs = _jobs(25)
    print("case serve mix", [j.natoms for j in jobs], flush=True)
    srv = Server(hbm_budget_bytes=64 << 30, max_slots=2, quantum=10)
    ids = [srv.submit(j) for j in jobs]
    assert srv.wait(300.0)
    res = {r["id"]: r for r in srv.poll()}
    srv.close()
    assert sorted(res) == sorted(ids)
    for jid, job in zip(ids, jobs):
        r = res[jid]
        assert r["status"] == 0, r["message"]
        assert r["nsteps"] == 25
        pos, vel, pairs, last = _alone(job)
        mpos, mvel = job.final_state()
        assert np.array_equal(pos, mpos) and np.array_equal(vel, mvel)
        assert r["pair_interactions"] == pairs
        assert r["epot"] == last[0] and r["temperature"] == last[2]
    # two slots: the third job waits for a free slot, then joins
    t = sorted((res[i]["t_admit"], res[i]["t_done"]) for i in ids)
    assert t[2][0] >= min(t[0][1], t[1][1]) - 1e-9


def test_admission_respects_hbm_budget():
    jobs = _jobs(10)
    est = [j.estimate_bytes() for j in jobs]
    print("case serve budget", est, flush=True)
    srv = Server(hbm_budget_bytes=max(est), max_slots=8, quantum=5)
    ids = [srv.submit(j) for j in jobs]
    assert srv.wait(300.0)
    res = {r["id"]: r for r in srv.poll()}
    srv.close()
    assert all(res[i]["status"] == 0 for i in ids)
    # at every admission the jobs resident together fit the budget
    iv = [(res[i]["t_admit"], res[i]["t_done"], res[i]["bytes"]) for i in ids]
    for t0, _, _ in iv:
        assert sum(b for a, d, b in iv if a <= t0 < d) <= max(est)
    big = Server(hbm_budget_bytes=min(est) // 2, max_slots=2, quantum=5)
    with pytest.raises(MdxError):
        big.submit(jobs[0])
    big.close()
