"""The batched configuration at full size: 2048 independent piles of
log-uniform 8-1024 boxes (the 10k pile's generator per scene, 429,041 boxes),
stepped as one resident batch (fixed-capacity frames replayed from a captured
HIP graph). Sixteen scenes spread over the size range are stepped alone by
the checker; every 10 frames their bodies and static-row lambdas must equal
the batch's slices bit for bit, over 40 frames (160 substeps, the bottom
boxes land). The reference steps one scene per Solver6DOF
(solver_6dof.py:177); a batch must give each scene exactly that result."""
import time

import numpy as np
import pytest

from avbd_amd import learned
from avbd_amd.batched import BatchedSolver6DOF
from oracle.oracle import oracle_solver_class

pytestmark = pytest.mark.gpu

N_SCENES = 2048


def _sizes():
    rng = np.random.default_rng(0)
    return np.round(np.exp(rng.uniform(np.log(8), np.log(1024), N_SCENES))).astype(int)


def test_2048_pile_batch_sampled_scenes_match_checker():
    sizes = _sizes()
    assert sizes.min() >= 8 and sizes.max() <= 1024
    t0 = time.perf_counter()
    parts = [learned.pile(int(n), seed=k) for k, n in enumerate(sizes)]
    batch = BatchedSolver6DOF(parts)
    batch.device = "cuda:0"   # the resident frame
    print(f"built {N_SCENES} scenes, {sum(len(p._x) for p in parts)} boxes "
          f"({time.perf_counter() - t0:.0f} s)", flush=True)
    order = np.argsort(sizes, kind="stable")
    pick = [int(k) for k in order[np.linspace(0, N_SCENES - 1, 16).round().astype(int)]]
    O = oracle_solver_class()
    alone = {k: learned.pile(int(sizes[k]), seed=k, solver_cls=O, device="cuda:0") for k in pick}
    n_rows = [len(p._rows) for p in parts]
    r0 = np.concatenate([[0], np.cumsum(n_rows)])
    batch._flush()
    for s in alone.values():
        s._flush()
    assert batch._resident_on(batch.device)
    for f in range(40):
        batch.step()
        for s in alone.values():
            s.step()
        if (f + 1) % 10:
            continue
        bb = batch._h.read_bodies(len(batch._x))
        lam_b = batch._h.read_rows(int(r0[-1]))["lambda"]
        for k, s in alone.items():
            sl = batch.scene_slice(k)
            ba = s._h.read_bodies(len(s._x))
            for key in ("x", "q", "v", "w"):
                assert np.array_equal(ba[key], bb[key][sl]), (f, k, key)
            assert np.array_equal(s._h.read_rows(n_rows[k])["lambda"],
                                  lam_b[r0[k]:r0[k + 1]]), (f, k)
        print(f"frame {f + 1}: 16 scenes bit-identical", flush=True)
    launches = batch._h.sync_count()[1]
    assert launches == 40 * batch.substeps - 1   # every substep after the first replays
    # the bottom boxes have landed: floor rows carry load in every sampled scene
    lam_b = batch._h.read_rows(int(r0[-1]))["lambda"]
    assert all((lam_b[r0[k]:r0[k + 1]] != 0).any() for k in pick)
