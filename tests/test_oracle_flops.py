"""The algorithmic FLOP counts of the solver roofline: one reference Newton
iteration per event, counted by the oracle's FLOP-counting build of
solveConstraints (pk.h:477-590), against the constants the bench uses."""
import numpy as np

import oracle
from palzilla_amd import caustics as CA
from palzilla_amd import scenes


def test_iteration_flops_match_counting_build():
    s = scenes.stack_scene(mirror_cells=6, layer="slab", layer_cells=8, photons=1 << 12, width=16)
    nodes, root, *_ = oracle.lbvh(s.tris, bits=32)
    recs, _ = oracle.photon_trace(s, nodes, root)
    paths = oracle.lookup_paths(recs, s.points)
    assert (paths["event"] >> 24).max() >= 3
    for K in (1, 2, 3):
        pk = paths[(paths["event"] >> 24) >= K][:20]
        for bits in range(1 << K):
            ev = (K << 24) | bits
            evals, steps = set(), set()
            for p in pk:
                q = p.copy()
                q["event"] = ev
                # eta = 1 keeps every refraction clear of total internal reflection
                e, st = oracle.count_flops(s, q, s.points[p["pixel"]]["p"], 1.0)
                evals.add(e)
                steps.add(st)
            want_eval, want_step = CA.iteration_flops(ev)
            assert evals == {want_eval}, CA.event_str(ev)
            # the step's triangle test (pk.h:574-587) short-circuits: u + v and
            # 1 + eps per vertex and eps * eps are counted only when reached
            assert want_step - (2 * K + 1) <= min(steps) and max(steps) <= want_step, CA.event_str(ev)


def test_flop_counter_counts_a_known_kernel():
    """The counting scalar on a hand-countable case: a flat mirror chain's
    evaluation is the same for every chain (no data-dependent arithmetic),
    and doubling the chain length more than doubles it (the 2K passes)."""
    e1, _ = CA.iteration_flops(1 << 24)
    e2, _ = CA.iteration_flops(2 << 24)
    e3, _ = CA.iteration_flops(3 << 24)
    assert e2 > 2 * e1 and e3 > 1.5 * e2
    # per pass and per refractive vertex: refraction_norm_free minus reflection
    assert CA.ITERATION_FLOPS_EVAL_PER_T[1] * 2 == CA.ITERATION_FLOPS_EVAL_PER_T[2]
    assert CA.ITERATION_FLOPS_EVAL_PER_T[1] * 3 == CA.ITERATION_FLOPS_EVAL_PER_T[3]
