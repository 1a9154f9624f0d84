"""Writes the golden fixtures of the parity tests (run from the repo root:
python tests/golden/make_golden.py).

Inputs come from the seeded scenario generators; expected outputs from the
CPU oracle (oracle/cd_oracle.c), whose own correctness is pinned by
tests/test_oracle_*.py.  The reference itself cannot run here (numba and
cell_lists are not importable), so it contributes no vectors."""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from crowddynamics_amd import examples as E  # noqa: E402
from oracle import oracle as O  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))


def scenario_cfg(s):
    nv = s.nav
    return O.Scenario(cell_size=s.cell_size, dt_min=s.dt_min, dt_max=s.dt_max, seed=s.seed,
                      segments=s.segments, nav_maps=nv['maps'] if nv else None,
                      nav_mode=nv['mode'] if nv else 0,
                      nav_origin=(nv['minx'], nv['miny']) if nv else (0.0, 0.0),
                      nav_step=nv['step'] if nv else 0.1)


def main():
    h = E.hallway()
    a1 = h.agents.copy()
    dt1 = O.run(a1, scenario_cfg(h), 1)
    a10 = h.agents.copy()
    dt10 = O.run(a10, scenario_cfg(h), 10)
    f0 = h.agents.copy()
    f0['force'] = 0.0
    fs, _ = O.agent_agent_scale(f0, h.cell_size)
    np.savez_compressed(os.path.join(HERE, 'hallway200.npz'), initial=h.agents,
                        step1=a1, step10=a10, dt1=dt1, dt10=dt10, segments=h.segments,
                        pair_force=f0['force'], pair_scale=fs, cell_size=h.cell_size)
    c = E.corridor(n=400, width=24.0, height=12.0, n_obstacles=6)
    t0 = c.agents.copy()
    t0['velocity'] += np.random.default_rng(3).normal(0, 0.3, t0['velocity'].shape)
    b = t0.copy()
    b['force'] = 0.0
    b['torque'] = 0.0
    fs, ts = O.agent_agent_scale(b, c.cell_size)
    t1 = t0.copy()
    dtt = O.run(t1, scenario_cfg(c), 1)
    np.savez_compressed(os.path.join(HERE, 'threecircle400.npz'), initial=t0, step1=t1,
                        dt1=dtt, segments=c.segments, pair_force=b['force'],
                        pair_torque=b['torque'], pair_scale=fs, torque_scale=ts,
                        cell_size=c.cell_size)
    print('golden fixtures written')


if __name__ == '__main__':
    main()
